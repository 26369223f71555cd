"""ctypes loader for the CPU oracle (oracle/liboracle_mr.so) used by the tests.

The oracle is test infrastructure; it is the checker, never the product."""

import ctypes as ct
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "numi-lab_amd"))

from metalrobo.native import (  # noqa: E402
    BODY_STATE_DTYPE,
    CONTACT_STATUS_DTYPE,
    MANIFOLD_HEADER_DTYPE,
    MANIFOLD_POINT_DTYPE,
    MR_MANIFOLD_POINT_CAPACITY,
    PAIR_RECORD_DTYPE,
    WORLD_STATUS_DTYPE,
    WorldModelDescC,
    WorldStepConfigC,
)

LIBRARY = os.path.join(ROOT, "oracle", "liboracle_mr.so")
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.isfile(LIBRARY):
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import build_oracle  # noqa: E402

            build_oracle.build(verbose=False)
        lib = ct.CDLL(LIBRARY)
        P = ct.c_void_p
        lib.mr_oracle_control_step.argtypes = [ct.POINTER(WorldModelDescC), ct.POINTER(WorldStepConfigC),
                                               ct.c_uint32, ct.c_uint32, ct.c_uint32] + [P] * 16
        lib.mr_oracle_control_step.restype = ct.c_int
        lib.mr_oracle_control_step_parallel.argtypes = [ct.POINTER(WorldModelDescC), ct.POINTER(WorldStepConfigC),
                                                        ct.c_uint32, ct.c_uint32, ct.c_uint32] + [P] * 10
        lib.mr_oracle_control_step_parallel.restype = ct.c_int
        lib.mr_oracle_forward_dynamics.argtypes = [ct.POINTER(WorldModelDescC), ct.POINTER(WorldStepConfigC),
                                                   ct.c_uint32, P, P, P, P]
        lib.mr_oracle_forward_dynamics.restype = ct.c_int
        lib.mr_oracle_point_response.argtypes = [ct.POINTER(WorldModelDescC), ct.POINTER(WorldStepConfigC),
                                                 ct.c_uint32, P, P, ct.c_uint32, P, P, P]
        lib.mr_oracle_point_response.restype = ct.c_int
        lib.mr_oracle_link_poses.argtypes = [ct.POINTER(WorldModelDescC), ct.POINTER(WorldStepConfigC),
                                             ct.c_uint32, P, P, P]
        lib.mr_oracle_link_poses.restype = ct.c_int
        lib.mr_oracle_collide.argtypes = [ct.POINTER(WorldModelDescC), ct.POINTER(WorldStepConfigC), ct.c_uint32,
                                          P, P, P, P, P, P, P]
        lib.mr_oracle_collide.restype = ct.c_int
        lib.mr_oracle_sincos.argtypes = [P, P, P, ct.c_uint32]
        lib.mr_oracle_sincos.restype = None
        lib.mr_oracle_abi_version.argtypes = []
        lib.mr_oracle_abi_version.restype = ct.c_uint32
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


class OracleState:
    """Host copy of one world's persistent state in the layout the oracle steps."""

    def __init__(self, model, config):
        E = model.environment_count
        self.model = model
        self.config = config
        self.q, self.v, self.scene = model.initial_state()
        cap = config.manifold_capacity
        self.mh = np.zeros((E, cap), MANIFOLD_HEADER_DTYPE)
        self.mp = np.zeros((E, cap, MR_MANIFOLD_POINT_CAPACITY), MANIFOLD_POINT_DTYPE)
        self.mcount = np.zeros(E, np.uint32)
        self.reset_q, self.reset_v, self.reset_scene = model.initial_state()

    def step(self, command, reset_mask=None, control_step=0, env_begin=0, env_end=None):
        """One control step; returns (status, contact status, acceleration, evidence pairs, counts)."""
        lib = load()
        E = self.model.environment_count
        env_end = E if env_end is None else env_end
        command = np.ascontiguousarray(command, np.float32)
        assert command.shape == (E, self.model.max_nv)
        status = np.zeros(E, WORLD_STATUS_DTYPE)
        cstatus = np.zeros(E, CONTACT_STATUS_DTYPE)
        acc = np.zeros((E, self.model.max_nv), np.float32)
        pairs = np.zeros((E, self.config.pair_capacity), PAIR_RECORD_DTYPE)
        counts = np.zeros(E, np.uint32)
        mask = None if reset_mask is None else np.ascontiguousarray(reset_mask, np.uint32)
        rc = lib.mr_oracle_control_step(ct.byref(self.model.desc), ct.byref(self.config), env_begin, env_end,
                                        control_step, _p(self.q), _p(self.v), _p(self.scene), _p(self.mh),
                                        _p(self.mp), _p(self.mcount), _p(command), _p(mask), _p(self.reset_q),
                                        _p(self.reset_v), _p(self.reset_scene), _p(status), _p(cstatus), _p(acc),
                                        _p(pairs), _p(counts))
        assert rc == 0
        return status, cstatus, acc, pairs, counts


def forward_dynamics(model, config, env, q, v, tau):
    lib = load()
    out = np.zeros(model.max_nv, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    tau = np.ascontiguousarray(tau, np.float32)
    code = lib.mr_oracle_forward_dynamics(ct.byref(model.desc), ct.byref(config), env, _p(q), _p(v), _p(tau),
                                          _p(out))
    return code, out


def point_response(model, config, env, q, v, link, point, force):
    lib = load()
    out = np.zeros(model.max_nv, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    point = np.ascontiguousarray(point, np.float32)
    force = np.ascontiguousarray(force, np.float32)
    code = lib.mr_oracle_point_response(ct.byref(model.desc), ct.byref(config), env, _p(q), _p(v), link, _p(point),
                                        _p(force), _p(out))
    return code, out


def link_poses(model, config, env, q, v):
    lib = load()
    n = int(model.models()[model.environment_model()[env]]["link_count"])
    out = np.zeros((n, 7), np.float32)
    q = np.ascontiguousarray(q, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    code = lib.mr_oracle_link_poses(ct.byref(model.desc), ct.byref(config), env, _p(q), _p(v), _p(out))
    return code, out


def collide(model, config, env, q, v, scene):
    """Per compiled pair: overlap flag, witness count, reals [8][10], features [8][2]."""
    lib = load()
    h = model.models()[model.environment_model()[env]]
    n = max(1, int(h["pair_count"]))
    overlap = np.zeros(n, np.uint32)
    counts = np.zeros(n, np.uint32)
    reals = np.zeros((n, 8, 10), np.float32)
    feats = np.zeros((n, 8, 2), np.uint32)
    q = np.ascontiguousarray(q, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    scene = np.ascontiguousarray(scene, BODY_STATE_DTYPE)
    rc = lib.mr_oracle_collide(ct.byref(model.desc), ct.byref(config), env, _p(q), _p(v), _p(scene), _p(overlap),
                               _p(counts), _p(reals), _p(feats))
    assert rc >= 0
    k = int(h["pair_count"])
    return overlap[:k], counts[:k], reals[:k], feats[:k]


def sincos(x):
    lib = load()
    x = np.ascontiguousarray(x, np.float32)
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    lib.mr_oracle_sincos(_p(x), _p(s), _p(c), x.size)
    return s, c


def exact_cone(inverse_mass, free_velocity, jacobian, target, regularization, friction_uv, warm=None,
               max_iterations=50000, tolerance=1.0e-11):
    """FP64 exact-cone solve (oracle/exact_cone.cpp). jacobian: [contacts*3][nv];
    friction_uv: [contacts][2]. Returns (code, impulses, velocity, stats dict)."""
    lib = load()
    if not hasattr(lib, "_exact_cone_bound"):
        P = ct.c_void_p
        lib.mr_oracle_exact_cone.argtypes = [ct.c_uint32, P, P, ct.c_uint32, P, P, P, P, P, ct.c_uint32,
                                             ct.c_double, P, P, P]
        lib.mr_oracle_exact_cone.restype = ct.c_int
        lib._exact_cone_bound = True
    M = np.ascontiguousarray(inverse_mass, np.float64)
    nv = M.shape[0]
    J = np.ascontiguousarray(jacobian, np.float64).reshape(-1, nv)
    n = J.shape[0]
    vf = np.ascontiguousarray(free_velocity, np.float64)
    tg = np.ascontiguousarray(target, np.float64).reshape(n)
    rg = np.ascontiguousarray(regularization, np.float64).reshape(n)
    mu = np.ascontiguousarray(friction_uv, np.float64).reshape(n // 3, 2)
    w = None if warm is None else np.ascontiguousarray(warm, np.float64).reshape(n)
    imp = np.zeros(n)
    vel = np.zeros(nv)
    stats = np.zeros(5)
    code = lib.mr_oracle_exact_cone(nv, _p(M), _p(vf), n // 3, _p(J), _p(tg), _p(rg), _p(mu),
                                    None if w is None else _p(w), max_iterations, tolerance, _p(imp), _p(vel),
                                    _p(stats))
    return code, imp, vel, {"iterations": int(stats[0]), "residual": stats[1], "violation": stats[2],
                            "used_warm": bool(stats[3]), "objective": stats[4]}


def exact_cone_program(hessian, linear, friction_uv, warm=None, max_iterations=200000, tolerance=1.0e-10):
    """FP64 exact-cone solve of a program given by H and l directly."""
    lib = load()
    P = ct.c_void_p
    lib.mr_oracle_exact_cone_program.argtypes = [ct.c_uint32, P, P, P, P, ct.c_uint32, ct.c_double, P, P]
    lib.mr_oracle_exact_cone_program.restype = ct.c_int
    H = np.ascontiguousarray(hessian, np.float64)
    n = H.shape[0]
    lin = np.ascontiguousarray(linear, np.float64)
    mu = np.ascontiguousarray(friction_uv, np.float64).reshape(n // 3, 2)
    w = None if warm is None else np.ascontiguousarray(warm, np.float64)
    imp = np.zeros(n)
    stats = np.zeros(5)
    code = lib.mr_oracle_exact_cone_program(n // 3, _p(H), _p(lin), _p(mu), None if w is None else _p(w),
                                            max_iterations, tolerance, _p(imp), _p(stats))
    return code, imp, {"iterations": int(stats[0]), "residual": stats[1], "violation": stats[2],
                       "used_warm": bool(stats[3]), "objective": stats[4]}


def contact_problem_capture(state, command, env, control_step=0, capacity=1 << 20):
    """One oracle control step that also returns environment `env`'s
    final-microstep contact problem: dict with J, R (rows x nv), v_free,
    target, reg, mu (contacts x 2), lambda_warm, lambda_final, v_final."""
    lib = load()
    lib.mr_oracle_set_problem_capture.argtypes = [ct.c_uint32, ct.c_void_p, ct.c_uint32]
    lib.mr_oracle_set_problem_capture.restype = None
    lib.mr_oracle_problem_capture_used.restype = ct.c_uint32
    buf = np.zeros(capacity, np.float32)
    lib.mr_oracle_set_problem_capture(env, _p(buf), capacity)
    try:
        out = state.step(command, control_step=control_step)
        used = int(lib.mr_oracle_problem_capture_used())
    finally:
        lib.mr_oracle_set_problem_capture(0xFFFFFFFF, None, 0)
    if used == 0:
        return out, None
    hdr = buf[:8].view(np.uint32)
    nv, nc, dyn = int(hdr[0]), int(hdr[1]), int(hdr[2])
    at = 8
    def take(k):
        nonlocal at
        a = buf[at:at + k].astype(np.float64)
        at += k
        return a
    prob = {"nv": nv, "contacts": nc, "dynamic_scene": bool(dyn)}
    prob["J"] = take(3 * nc * nv).reshape(3 * nc, nv)
    prob["R"] = take(3 * nc * nv).reshape(3 * nc, nv)
    prob["v_free"] = take(nv)
    prob["target"] = take(3 * nc)
    prob["reg"] = take(3 * nc)
    prob["mu"] = take(2 * nc).reshape(nc, 2)
    prob["lambda_warm"] = take(3 * nc)
    ns = int(hdr[3])
    geo = take(16 * nc).reshape(nc, 16)
    prob["dyn"] = geo[:, 0:2] > 0.5
    prob["scene_index"] = geo[:, 2:4].astype(np.int64)
    prob["point"] = geo[:, 4:7]
    prob["dir"] = geo[:, 7:16].reshape(nc, 3, 3)
    sc = take(19 * ns).reshape(ns, 19)
    prob["scene_pos"] = sc[:, 0:3]
    prob["scene_inv_mass"] = sc[:, 3]
    prob["scene_inv_inertia"] = sc[:, 4:13].reshape(ns, 3, 3)
    prob["scene_twist_free"] = sc[:, 13:19]
    prob["lambda_final"] = take(3 * nc)
    prob["v_final"] = take(nv)
    prob["scene_twist_final"] = take(6 * ns).reshape(ns, 6)
    assert at == used, (at, used)
    return out, prob


def contact_program(prob):
    """The captured problem as one dense cone program over the articulation
    DoFs and the scene bodies' twists: returns (H, l, apply) where
    apply(lambda) gives (v, twists) after the impulses. J rows: +d on body B,
    -d on body A (linear d, angular r x d on scene bodies)."""
    nc, ns = prob["contacts"], len(prob["scene_inv_mass"])
    n = 3 * nc
    Js = np.zeros((n, 6 * ns))
    for c in range(nc):
        for side, sign in ((0, -1.0), (1, 1.0)):
            if not prob["dyn"][c, side]:
                continue
            b = int(prob["scene_index"][c, side])
            lever = prob["point"][c] - prob["scene_pos"][b]
            for r in range(3):
                d = prob["dir"][c, r]
                Js[3 * c + r, 6 * b:6 * b + 3] = sign * d
                Js[3 * c + r, 6 * b + 3:6 * b + 6] = sign * np.cross(lever, d)
    Minv = np.zeros((6 * ns, 6 * ns))
    for b in range(ns):
        Minv[6 * b:6 * b + 3, 6 * b:6 * b + 3] = prob["scene_inv_mass"][b] * np.eye(3)
        Minv[6 * b + 3:6 * b + 6, 6 * b + 3:6 * b + 6] = prob["scene_inv_inertia"][b]
    J, R = prob["J"], prob["R"]
    H = J @ R.T + Js @ Minv @ Js.T + np.diag(prob["reg"])
    H = 0.5 * (H + H.T)
    tw0 = prob["scene_twist_free"].reshape(-1)
    lin = J @ prob["v_free"] + Js @ tw0 - prob["target"]

    def apply(lam):
        return prob["v_free"] + R.T @ lam, (tw0 + Minv @ Js.T @ lam).reshape(ns, 6)

    return H, lin, apply


# ------------------------------------------------------------- task layer

def _task_lib():
    lib = load()
    if not getattr(lib, "_task_bound", False):
        from metalrobo.native import LocomotionTaskProgramC
        P = ct.c_void_p
        lib.mr_oracle_mix64.argtypes = [ct.c_uint64]
        lib.mr_oracle_mix64.restype = ct.c_uint64
        lib.mr_oracle_task_random_unit.argtypes = [ct.c_uint64, ct.c_uint32, ct.c_uint32, ct.c_uint32, ct.c_uint32]
        lib.mr_oracle_task_random_unit.restype = ct.c_float
        prog = ct.POINTER(LocomotionTaskProgramC)
        lib.mr_oracle_task_apply_actions.argtypes = [prog, ct.c_uint32, ct.c_uint32, ct.c_float, P, P, P, P]
        lib.mr_oracle_task_apply_actions.restype = None
        lib.mr_oracle_task_complete.argtypes = ([prog, ct.c_uint32, ct.c_uint32, ct.c_uint32, ct.c_float] + [P] * 8
                                                + [ct.c_uint64])
        lib.mr_oracle_task_complete.restype = None
        lib.mr_oracle_task_observe.argtypes = [prog, ct.c_uint32, ct.c_uint32, ct.c_uint32, ct.c_float] + [P] * 9
        lib.mr_oracle_task_observe.restype = None
        lib.mr_oracle_policy_dense.argtypes = [P, P, P, P, ct.c_float, P, ct.c_uint32, P, ct.c_uint32, ct.c_uint32,
                                               ct.c_uint32, ct.c_uint32, ct.c_uint32]
        lib.mr_oracle_policy_dense.restype = None
        lib.mr_oracle_policy_sample.argtypes = [P, ct.c_uint32, P, P, P, ct.c_float, ct.c_uint64, ct.c_uint64, P,
                                                ct.c_uint32, ct.c_uint32, ct.c_uint32, P, P, P]
        lib.mr_oracle_policy_sample.restype = None
        lib._task_bound = True
    return lib


class OracleTask:
    """The task layer's state and the oracle restatement of its kernels, stepping beside an
    OracleState for the physics (apply_actions -> physics -> complete -> observe)."""

    def __init__(self, model, config, program, physics=None):
        from metalrobo.native import (TASK_ACTION_DTYPE, TASK_OBSERVATION_DTYPE, TASK_REWARD_DTYPE,
                                      TASK_STATE_DTYPE, TASK_TERMINATION_DTYPE, TASK_TRANSITION_DTYPE,
                                      task_program_struct)
        self.lib = _task_lib()
        self.model, self.config = model, config
        E, nq, nv = model.environment_count, model.max_nq, model.max_nv
        self.arrays = {k: np.ascontiguousarray(program[k], dt) for k, dt in (
            ("actions", TASK_ACTION_DTYPE), ("observations", TASK_OBSERVATION_DTYPE),
            ("rewards", TASK_REWARD_DTYPE), ("terminations", TASK_TERMINATION_DTYPE))}
        self.arrays["default_q"] = np.ascontiguousarray(program["default_q"], np.float32)
        self.prog = task_program_struct(program, self.arrays)
        self.A, self.O, self.H = len(self.arrays["actions"]), len(self.arrays["observations"]), int(program["history_length"])
        self.slots = int(program["delay_slots"])
        self.dt = np.float32(config.control_timestep)
        self.physics = physics if physics is not None else OracleState(model, config)
        ph = self.physics
        ph.reset_q = np.tile(self.arrays["default_q"], (E, 1)).astype(np.float32)
        ph.reset_v = np.zeros((E, nv), np.float32)
        ph.reset_scene = model.initial_state()[2].copy()
        self.states = np.zeros(E, TASK_STATE_DTYPE)
        self.policy_revision = 0  # stamped on every transition
        self.history = np.zeros((E, self.slots, self.A), np.float32)
        self.previous = np.zeros((E, 2, self.A), np.float32)
        self.observations = np.zeros((E, self.H, self.O), np.float32)
        self.transitions = np.zeros(E, TASK_TRANSITION_DTYPE)
        self.targets = np.zeros((E, nv), np.float32)
        self.mask = np.zeros(E, np.uint32)
        self.dofs = np.ascontiguousarray(model.dofs()[model.models()[0]["dof_offset"]:])
        self.observe()

    def apply_actions(self, actions):
        E, nv = self.model.environment_count, self.model.max_nv
        actions = np.ascontiguousarray(actions, np.float32)
        self.lib.mr_oracle_task_apply_actions(ct.byref(self.prog), E, nv, self.dt, _p(self.states), _p(self.history),
                                              _p(actions), _p(self.targets))

    def complete(self, status_codes):
        E, nq, nv = self.model.environment_count, self.model.max_nq, self.model.max_nv
        codes = np.ascontiguousarray(status_codes, np.uint32)
        ph = self.physics
        self.lib.mr_oracle_task_complete(ct.byref(self.prog), E, nq, nv, self.dt, _p(ph.q), _p(ph.v), _p(self.dofs),
                                         _p(codes), _p(self.states), _p(self.history), _p(self.previous),
                                         _p(self.transitions), self.policy_revision)

    def observe(self):
        E, nq, nv = self.model.environment_count, self.model.max_nq, self.model.max_nv
        ph = self.physics
        self.lib.mr_oracle_task_observe(ct.byref(self.prog), E, nq, nv, self.dt, _p(ph.q), _p(ph.v), _p(ph.reset_q),
                                        _p(ph.reset_v), _p(self.states), _p(self.history), _p(self.previous),
                                        _p(self.observations), _p(self.mask))

    def step(self, actions, control_step=0):
        self.apply_actions(actions)
        status, *_ = self.physics.step(self.targets, reset_mask=self.mask.copy(), control_step=control_step)
        self.complete(status["code"])
        self.observe()
        return status


def policy_dense(weights, bias, inp, out_width, activation, mean=None, inverse_std=None, clip=0.0):
    lib = _task_lib()
    inp = np.ascontiguousarray(inp, np.float32)
    E, in_w = inp.shape
    out = np.zeros((E, out_width), np.float32)
    lib.mr_oracle_policy_dense(_p(np.ascontiguousarray(weights, np.float32)), _p(np.ascontiguousarray(bias, np.float32)),
                               None if mean is None else _p(np.ascontiguousarray(mean, np.float32)),
                               None if inverse_std is None else _p(np.ascontiguousarray(inverse_std, np.float32)),
                               clip, _p(inp), in_w, _p(out), out_width, E, int(weights.shape[1]), out_width, activation)
    return out


def policy_sample(mean, action_bias, action_scale, log_std, action_clip, seed, revision, states, control_step):
    lib = _task_lib()
    mean = np.ascontiguousarray(mean, np.float32)
    E, A = mean.shape
    acts = np.zeros((E, A), np.float32)
    lat = np.zeros((E, A), np.float32)
    lp = np.zeros(E, np.float32)
    keep = [np.ascontiguousarray(x, np.float32) for x in (action_bias, action_scale)]
    ls = None if log_std is None else np.ascontiguousarray(log_std, np.float32)
    lib.mr_oracle_policy_sample(_p(mean), A, _p(keep[0]), _p(keep[1]), None if ls is None else _p(ls), action_clip,
                                seed, revision, _p(np.ascontiguousarray(states)), E, A, control_step, _p(acts), _p(lat),
                                _p(lp))
    return acts, lat, lp
