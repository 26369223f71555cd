"""Helpers that build small custom worlds through mr_world_model_create, for
known-answer and invariant tests (records follow include/mr_world_types.h)."""

import ctypes as ct

import numpy as np

import metalrobo as mr
from metalrobo import native as nat

INVALID = 0xFFFFFFFF
Y_TO_Z = np.array([np.sqrt(0.5), 0.0, 0.0, np.sqrt(0.5)], np.float32)


def body(mass, inertia, motion=2, parent=INVALID, joint_type=15, q_index=INVALID, v_index=INVALID,
         axis=(1, 0, 0), parent_anchor=(0, 0, 0), child_anchor=(0, 0, 0), damping=(0, 0, 0, 0)):
    b = np.zeros(1, nat.BODY_DTYPE)[0]
    b["parent"] = parent
    b["joint_type"] = joint_type
    b["q_index"] = q_index
    b["v_index"] = v_index
    b["motion_type"] = motion
    dyn = motion == 2 and mass > 0
    b["mass"] = [mass, 1.0 / mass if dyn else 0.0, 0, 0]
    inertia = np.asarray(inertia, np.float64)
    if inertia.ndim == 1:
        inertia = np.diag(inertia)
    b["inertia"][:, :3] = inertia
    b["inverse_inertia"][:, :3] = np.linalg.inv(inertia) if dyn else 0.0
    b["axis"][:3] = axis
    b["parent_anchor"][:3] = parent_anchor
    b["child_anchor"][:3] = child_anchor
    b["parent_rotation"] = [0, 0, 0, 1]
    b["child_rotation"] = [0, 0, 0, 1]
    b["damping"] = damping
    b["damping_scale"] = [1, 1, 0, 0]
    return b


def shape(body_index, shape_type, dims, pos=(0, 0, 0), rot=(0, 0, 0, 1), contact_offset=0.01, material=0,
          geometry=0, flags=0):
    s = np.zeros(1, nat.SHAPE_DTYPE)[0]
    s["body"] = body_index
    s["shape_type"] = shape_type
    s["material"] = material
    s["collision_group"] = 1
    s["collision_mask"] = 0xFFFFFFFF
    s["local_position"][:3] = pos
    s["local_rotation"] = rot
    s["dimensions"][:len(dims)] = dims
    s["contact_rest_radius"] = [contact_offset, 0, 0, 0]
    s["geometry"] = geometry
    s["flags"] = flags
    return s


def geometry_tables(geoms):
    """(geometries, vertices, triangles) arrays from [(kind, vertices [n][3], triangles [m][3] or None)];
    triangle feature ids are their indices."""
    g = np.zeros(len(geoms), nat.GEOMETRY_DTYPE)
    verts, tris = [], []
    for k, (kind, vs, ts) in enumerate(geoms):
        g[k]["kind"] = kind
        g[k]["vertex_offset"] = len(verts)
        g[k]["vertex_count"] = len(vs)
        verts += [[*v, 0.0] for v in vs]
        if ts is not None:
            g[k]["triangle_offset"] = len(tris)
            g[k]["triangle_count"] = len(ts)
            tris += [(tuple(t), i) for i, t in enumerate(ts)]
    vt = np.array(verts, np.float32).reshape(-1, 4)
    tr = np.zeros(len(tris), nat.TRIANGLE_DTYPE)
    for i, (t, f) in enumerate(tris):
        tr[i]["vertices"] = t
        tr[i]["feature"] = f
    return g, vt, tr


def dof(body_index, q_index, flags=0, limits=(0, 0, 0, 0), drive=(0, 0, 0, 0)):
    d = np.zeros(1, nat.DOF_DTYPE)[0]
    d["body"] = body_index
    d["q_index"] = q_index
    d["flags"] = flags
    d["limits"] = limits
    d["drive"] = drive
    return d


def actuator(local_v, torque_constant=0.0, current_limit=0.0, no_load_speed=0.0, efficiency=0.0, active=True):
    """MRActuatorProfileC: stall torque = torque constant x current limit."""
    a = np.zeros(1, nat.ACTUATOR_DTYPE)[0]
    if active:
        a["motor_and_speed"] = [torque_constant, current_limit, no_load_speed, efficiency]
        a["transmission_and_envelope"] = [0, 0, np.float32(torque_constant) * np.float32(current_limit), 0]
    a["identity"] = [local_v, nat.MR_ACTUATOR_PROFILE_ACTIVE if active else 0, 0, 0]
    return a


def pair_class(a, b):
    lo, hi = min(a, b), max(a, b)
    table = {(0, 0): 1, (0, 3): 2, (1, 3): 3, (2, 3): 4, (0, 1): 6, (1, 1): 7, (0, 2): 8, (1, 2): 9, (2, 2): 10}
    if hi == 7 and lo <= 2:
        return 12
    if (hi == 5 and lo <= 3) or (lo, hi) == (5, 5):
        return 11  # convex hull
    if hi == 6 and (lo <= 2 or lo == 5):
        return 13  # triangle mesh
    return table.get((lo, hi), 0)


def compile_pairs(bodies, shapes, link_count, exclusion_distance=1):
    def depth(i):
        d = 0
        while bodies[i]["parent"] != INVALID:
            i = int(bodies[i]["parent"])
            d += 1
        return d

    def dist(a, b):
        if a >= link_count or b >= link_count:
            return 1 << 30
        da, db, n = depth(a), depth(b), 0
        while da > db:
            a, da, n = int(bodies[a]["parent"]), da - 1, n + 1
        while db > da:
            b, db, n = int(bodies[b]["parent"]), db - 1, n + 1
        while a != b:
            a, b, n = int(bodies[a]["parent"]), int(bodies[b]["parent"]), n + 2
        return n

    out = []
    for a in range(len(shapes)):
        for b in range(a + 1, len(shapes)):
            sa, sb = shapes[a], shapes[b]
            if sa["body"] == sb["body"]:
                continue
            if (int(sa["collision_group"]) & int(sb["collision_mask"])) == 0 or \
                    (int(sb["collision_group"]) & int(sa["collision_mask"])) == 0:
                continue
            if bodies[sa["body"]]["motion_type"] != 2 and bodies[sb["body"]]["motion_type"] != 2:
                continue
            if dist(int(sa["body"]), int(sb["body"])) <= exclusion_distance:
                continue
            cls = pair_class(int(sa["shape_type"]), int(sb["shape_type"]))
            if cls:
                out.append((a, b, cls, 0))
    return np.array(out, nat.PAIR_DTYPE) if out else np.zeros(0, nat.PAIR_DTYPE)


class CustomWorld:
    """One model block shared by `envs` environments."""

    def __init__(self, bodies, dofs, shapes, link_count, root_type, nq, nv, materials=None, envs=1,
                 gravity=(0, 0, -9.81), microstep=0.005, exclusion_distance=1, heights=None, hf_origin=None,
                 hf_spacing=None, actuators=None, geometries=None):
        self.bodies = np.array(bodies, nat.BODY_DTYPE)
        self.dofs = np.array(dofs, nat.DOF_DTYPE) if len(dofs) else np.zeros(1, nat.DOF_DTYPE)
        self.shapes = np.array(shapes, nat.SHAPE_DTYPE) if len(shapes) else np.zeros(1, nat.SHAPE_DTYPE)
        nshapes = len(shapes)
        if materials is None:
            materials = [((0.8, 0.8, 0.8, 0.8), (0.0, 0.5, 0.0, 0.0))]
        self.materials = np.array(materials, nat.MATERIAL_DTYPE)
        self.pairs = compile_pairs(self.bodies, self.shapes[:nshapes], link_count, exclusion_distance)
        if len(self.pairs) == 0:
            self.pairs = np.zeros(1, nat.PAIR_DTYPE)
            npairs = 0
        else:
            npairs = len(self.pairs)
        hdr = np.zeros(1, nat.MODEL_HEADER_DTYPE)
        hdr["body_count"] = len(bodies)
        hdr["link_count"] = link_count
        hdr["root_type"] = root_type
        hdr["nq"] = nq
        hdr["nv"] = nv
        hdr["shape_count"] = nshapes
        hdr["pair_count"] = npairs
        hdr["scene_body_count"] = len(bodies) - link_count
        hdr["material_count"] = len(self.materials)
        self.header = hdr
        self.env_model = np.zeros(envs, np.uint32)
        self.heights = np.zeros(1, np.float32) if heights is None else np.ascontiguousarray(heights, np.float32)
        desc = nat.WorldModelDescC()
        desc.model_count = 1
        desc.environment_count = envs
        desc.max_nq = nq
        desc.max_nv = nv
        desc.max_scene_bodies = len(bodies) - link_count
        desc.gravity = nat._F4(*gravity, 0.0)
        desc.models = self.header.ctypes.data
        desc.environment_model = self.env_model.ctypes.data
        desc.bodies = self.bodies.ctypes.data
        desc.dofs = self.dofs.ctypes.data
        desc.shapes = self.shapes.ctypes.data
        desc.materials = self.materials.ctypes.data
        desc.pairs = self.pairs.ctypes.data
        desc.body_total = len(self.bodies)
        desc.dof_total = len(self.dofs)
        desc.shape_total = len(self.shapes)
        desc.material_total = len(self.materials)
        desc.pair_total = len(self.pairs)
        if heights is not None:
            ny, nx = self.heights.shape
            desc.heightfield_vertex_count = nx * ny
            desc.heightfield.nx = nx
            desc.heightfield.ny = ny
            desc.heightfield.present = 1
            desc.heightfield.origin_spacing = nat._F4(hf_origin[0], hf_origin[1], hf_spacing, 0)
            lo = [hf_origin[0], hf_origin[1], float(self.heights.min())]
            hi = [hf_origin[0] + hf_spacing * (nx - 1), hf_origin[1] + hf_spacing * (ny - 1),
                  float(self.heights.max())]
            desc.heightfield.bounds_lower = nat._F4(*lo, 0)
            desc.heightfield.bounds_upper = nat._F4(*hi, 0)
            desc.heightfield_heights = self.heights.ctypes.data
        if actuators is not None:
            self.actuators = np.array(actuators, nat.ACTUATOR_DTYPE)
            desc.actuators = self.actuators.ctypes.data
            desc.actuator_total = len(self.actuators)
        if geometries is not None:
            # a list of (kind, vertices, triangles), or the three tables already built
            tables = geometries if isinstance(geometries, tuple) else geometry_tables(geometries)
            self.geometries, self.geometry_vertices, self.triangles = tables
            desc.geometries = self.geometries.ctypes.data
            desc.geometry_total = len(self.geometries)
            desc.geometry_vertices = self.geometry_vertices.ctypes.data
            desc.geometry_vertex_total = len(self.geometry_vertices)
            if len(self.triangles):
                desc.triangles = self.triangles.ctypes.data
                desc.triangle_total = len(self.triangles)
        self._desc = desc  # kept (with the arrays it points into) for tests that alter a copy
        b = mr.load()
        h = b.lib.mr_world_model_create(ct.byref(desc), microstep)
        if not h:
            raise mr.MetalRoboError(b.last_error())
        self.model = mr.WorldModel(h, b)
        self.envs = envs
        self.nq = nq
        self.nv = nv

    def config(self, **overrides):
        cfg = nat.WorldStepConfigC()
        cfg.control_timestep = 0.02
        cfg.physics_substeps = 4
        cfg.velocity_iterations = 4
        cfg.final_velocity_iterations = 2
        cfg.actuation_mode = 0
        cfg.warm_start = 1
        cfg.apply_body_damping = 1
        cfg.pair_capacity = 64
        cfg.manifold_capacity = 32
        cfg.constraint_capacity = 64
        cfg.manifold_breaking_separation = 0.02
        cfg.manifold_breaking_tangential = 0.02
        cfg.manifold_merge_distance = 0.002
        cfg.manifold_normal_cosine = 0.95
        cfg.contact_slop = 0.001
        cfg.max_depenetration_velocity = 1.0
        cfg.residual_tolerance = 0.0
        cfg.max_iterations = 64
        for k, v in overrides.items():
            setattr(cfg, k, v)
        return cfg


def scene_state(pos, rot=(0, 0, 0, 1), lin=(0, 0, 0), ang=(0, 0, 0), inv_mass=0.0, motion=2):
    s = np.zeros(1, nat.BODY_STATE_DTYPE)[0]
    s["position"] = [*pos, 1.0]
    s["orientation"] = rot
    s["linear_velocity_inv_mass"] = [*lin, inv_mass]
    s["angular_velocity"][:3] = ang
    s["flags_indices"] = [motion, INVALID, 0, 0]
    return s
