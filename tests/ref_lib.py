"""ctypes loader for oracle/_ref/libmr_ref.so: the reference's own FP64 CPU
oracle (collision, articulated dynamics, exact-cone solve) compiled from its
sources by oracle/build_ref.py, behind the C harness oracle/ref_harness.cpp.

Test infrastructure only. The library exists where the reference sources were
present at build time; tests that need it skip otherwise."""

import ctypes as ct
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBRARY = os.path.join(ROOT, "oracle", "_ref", "libmr_ref.so")
INVALID = 0xFFFFFFFF
_lib = None


class RefLink(ct.Structure):
    _fields_ = [("parent", ct.c_uint32), ("joint_type", ct.c_uint32), ("q_index", ct.c_uint32),
                ("v_index", ct.c_uint32), ("mass", ct.c_float), ("inertia", ct.c_float * 9),
                ("axis", ct.c_float * 3), ("parent_anchor", ct.c_float * 3), ("child_anchor", ct.c_float * 3),
                ("parent_rotation", ct.c_float * 4), ("child_rotation", ct.c_float * 4),
                ("damping", ct.c_float * 2)]


def available() -> bool:
    return os.path.isfile(LIBRARY)


def load():
    global _lib
    if _lib is None:
        lib = ct.CDLL(LIBRARY)
        P = ct.c_void_p
        lib.mr_ref_forward_dynamics.argtypes = [ct.c_uint32, ct.c_uint32, P, ct.c_uint32, ct.c_uint32, P, P,
                                                ct.c_double, ct.c_uint32, P, P, P, P, ct.c_char_p, ct.c_size_t]
        lib.mr_ref_forward_dynamics.restype = ct.c_int
        lib.mr_ref_collide.argtypes = [ct.c_uint32, P, ct.c_uint32, P, ct.c_uint32, P, ct.c_uint32, ct.c_uint32,
                                       P, P, P, P, P, ct.c_char_p, ct.c_size_t]
        lib.mr_ref_collide.restype = ct.c_int
        lib.mr_ref_conic_solve.argtypes = [ct.c_uint32, P, P, ct.c_uint32, P, P, P, P, P, ct.c_uint32,
                                           ct.c_double, P, P, P, ct.c_char_p, ct.c_size_t]
        lib.mr_ref_conic_solve.restype = ct.c_int
        lib.mr_ref_pgs_solve.argtypes = [ct.c_uint32, P, ct.c_uint32, P, ct.c_double, ct.c_double, ct.c_double,
                                         ct.c_double, ct.c_uint32, P, ct.c_char_p, ct.c_size_t]
        lib.mr_ref_pgs_solve.restype = ct.c_int
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def tree_links(header, bodies):
    """RefLink records of one model block's links (include/mr_world_types.h MRBodyC)."""
    n = int(header["link_count"])
    b = bodies[int(header["body_offset"]):int(header["body_offset"]) + n]
    links = (RefLink * n)()
    for i in range(n):
        r = b[i]
        L = links[i]
        L.parent = int(r["parent"]) if i > 0 else INVALID
        L.joint_type = int(r["joint_type"])
        L.q_index = int(r["q_index"])
        L.v_index = int(r["v_index"])
        L.mass = float(r["mass"][0])
        inertia = np.asarray(r["inertia"])[:, :3].reshape(-1)
        L.inertia[:] = [float(x) for x in inertia]
        L.axis[:] = [float(x) for x in r["axis"][:3]]
        L.parent_anchor[:] = [float(x) for x in r["parent_anchor"][:3]]
        L.child_anchor[:] = [float(x) for x in r["child_anchor"][:3]]
        L.parent_rotation[:] = [float(x) for x in r["parent_rotation"]]
        L.child_rotation[:] = [float(x) for x in r["child_rotation"]]
        L.damping[:] = [float(r["damping"][0]), float(r["damping"][1])]
    return links


def forward_dynamics(header, bodies, dofs, q, v, tau, gravity=(0.0, 0.0, -9.81), timestep=0.005, damping=True):
    """FP64 generalized accelerations from the reference's computeArticulatedForwardDynamics."""
    lib = load()
    links = tree_links(header, bodies)
    nq, nv = int(header["nq"]), int(header["nv"])
    d = dofs[int(header["dof_offset"]):int(header["dof_offset"]) + nv]
    arm = np.ascontiguousarray([float(x["drive"][2]) for x in d], np.float32)
    grav = np.ascontiguousarray(gravity, np.float32)
    q = np.ascontiguousarray(q, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    tau = np.ascontiguousarray(tau, np.float64)
    out = np.zeros(nv, np.float64)
    err = ct.create_string_buffer(512)
    rc = lib.mr_ref_forward_dynamics(int(header["root_type"]), len(links), ct.addressof(links), nq, nv, _p(arm),
                                     _p(grav), timestep, 1 if damping else 0, _p(q), _p(v), _p(tau), _p(out), err,
                                     512)
    return rc, out, err.value.decode()


def collide(shapes, body_states, exclusions, pair_capacity=4096, raw_capacity=32768):
    """Reference FP64 collision of one environment from an empty manifold cache.
    Returns (rc, pairs [n][2], raw reals [m][10], raw u32 [m][3] = (feature A, feature B, pair index), err)."""
    lib = load()
    shapes = np.ascontiguousarray(shapes)
    assert shapes.dtype.itemsize == 96
    body_states = np.ascontiguousarray(body_states)
    assert body_states.dtype.itemsize == 128
    ex = np.ascontiguousarray(np.asarray(exclusions, np.uint32).reshape(-1, 2))
    pairs = np.zeros((pair_capacity, 2), np.uint32)
    raw = np.zeros((raw_capacity, 10), np.float32)
    raw_u = np.zeros((raw_capacity, 3), np.uint32)
    npairs = np.zeros(1, np.uint32)
    nraw = np.zeros(1, np.uint32)
    err = ct.create_string_buffer(512)
    rc = lib.mr_ref_collide(len(shapes), _p(shapes), len(body_states), _p(body_states), len(ex), _p(ex),
                            pair_capacity, raw_capacity, _p(pairs), _p(npairs), _p(raw), _p(raw_u), _p(nraw), err,
                            512)
    n, m = int(npairs[0]), int(nraw[0])
    return rc, pairs[:n], raw[:m], raw_u[:m], err.value.decode()


def conic_solve(inverse_mass, free_velocity, jacobian, target, regularization, friction, warm=None,
                max_iterations=200000, tolerance=1.0e-10):
    """Reference exact circular-cone solve. jacobian [contacts][3][nv]. Returns
    (rc, velocity, impulses [contacts][3], info = (iterations, objective, residual, cone violation), err)."""
    lib = load()
    M = np.ascontiguousarray(inverse_mass, np.float64)
    nv = M.shape[0]
    vf = np.ascontiguousarray(free_velocity, np.float64)
    J = np.ascontiguousarray(jacobian, np.float64)
    k = J.shape[0]
    tg = np.ascontiguousarray(target, np.float64).reshape(k, 3)
    rg = np.ascontiguousarray(regularization, np.float64).reshape(k, 3)
    mu = np.ascontiguousarray(friction, np.float64).reshape(k)
    w = None if warm is None else np.ascontiguousarray(warm, np.float64).reshape(k, 3)
    vel = np.zeros(nv, np.float64)
    lam = np.zeros((k, 3), np.float64)
    info = np.zeros(4, np.float64)
    err = ct.create_string_buffer(512)
    rc = lib.mr_ref_conic_solve(nv, _p(M), _p(vf), k, _p(J), _p(tg), _p(rg), _p(w), _p(mu), max_iterations,
                                tolerance, _p(vel), _p(lam), _p(info), err, 512)
    return rc, vel, lam, info, err.value.decode()


# The reference's free-body state and contact constraint records (engine_types.h:418-434, 1538-1563).
PGS_BODY_DTYPE = np.dtype([("position", "<f4", 4), ("orientation", "<f4", 4), ("linear_velocity_inverse_mass", "<f4", 4),
                           ("angular_velocity", "<f4", 4), ("inverse_inertia_world", "<f4", (3, 4)),
                           ("flags_and_indices", "<u4", 4)])
PGS_CONTACT_DTYPE = np.dtype([("body_a", "<u4"), ("body_b", "<u4"), ("flags", "<u4"), ("island", "<u4"),
                              ("pair_key", "<u8"), ("feature_key", "<u8"), ("point_and_separation", "<f4", 4),
                              ("normal", "<f4", 4), ("tangent", "<f4", 4), ("friction", "<f4", 4),
                              ("response", "<f4", 4), ("target_velocity_and_pre_solve_normal", "<f4", 4),
                              ("impulses", "<f4", 4)])
assert PGS_BODY_DTYPE.itemsize == 128 and PGS_CONTACT_DTYPE.itemsize == 144


def pgs_solve(bodies, contacts, timestep, error_reduction, penetration_slop, max_depenetration_velocity,
              iterations):
    """Reference projected Gauss-Seidel over free bodies, in place on copies.
    Returns (rc, bodies, contacts, info = (iterations, last maximum impulse change,
    normal residual, cone violation), err)."""
    lib = load()
    b = np.array(bodies, dtype=PGS_BODY_DTYPE, copy=True)
    c = np.array(contacts, dtype=PGS_CONTACT_DTYPE, copy=True)
    info = np.zeros(4, np.float64)
    err = ct.create_string_buffer(512)
    rc = lib.mr_ref_pgs_solve(len(b), _p(b), len(c), _p(c), timestep, error_reduction, penetration_slop,
                              max_depenetration_velocity, iterations, _p(info), err, 512)
    return rc, b, c, info, err.value.decode()
