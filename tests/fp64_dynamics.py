"""Independent FP64 rigid-tree dynamics used to pin the oracle's ABA.

Formulation (not the articulated-body algorithm): geometric Jacobians of every
link COM give the joint-space mass matrix M = sum m J_v^T J_v + J_w^T I_w J_w
plus the effective armature; the velocity-product and gravity bias h comes
from a recursive Newton-Euler pass with zero joint acceleration. Forward
dynamics is qdd = M^{-1} (tau - h), as in the reference's FP64 CRBA + Cholesky
oracle (src/core/ArticulatedDynamics.cpp:2241-2330). Frames follow the
engine convention: world-aligned, link poses at the COM, root linear velocity
at the root COM, v = [v_root, w_root, joint rates].
"""

import numpy as np

ROOT_FLOATING = 1


def qmul(l, r):
    lx, ly, lz, lw = l
    rx, ry, rz, rw = r
    return np.array([lw * rx + lx * rw + ly * rz - lz * ry, lw * ry - lx * rz + ly * rw + lz * rx,
                     lw * rz + lx * ry - ly * rx + lz * rw, lw * rw - lx * rx - ly * ry - lz * rz])


def qconj(q):
    return np.array([-q[0], -q[1], -q[2], q[3]])


def qrot(q, v):
    u = q[:3]
    t = 2.0 * np.cross(u, v)
    return v + q[3] * t + np.cross(u, t)


def qmat(q):
    return np.stack([qrot(q, e) for e in np.eye(3)], axis=1)


def axis_angle(axis, angle):
    a = axis / np.linalg.norm(axis)
    return np.concatenate([a * np.sin(0.5 * angle), [np.cos(0.5 * angle)]])


class TreeModel:
    """FP64 view of one model block (see include/mr_world_types.h)."""

    def __init__(self, header, bodies, dofs, gravity):
        self.n = int(header["link_count"])
        self.nv = int(header["nv"])
        self.nq = int(header["nq"])
        self.floating = int(header["root_type"]) == ROOT_FLOATING
        b = bodies[int(header["body_offset"]):int(header["body_offset"]) + self.n]
        self.bodies = b
        self.dofs = dofs[int(header["dof_offset"]):int(header["dof_offset"]) + self.nv]
        self.gravity = np.asarray(gravity[:3], np.float64)

    def kinematics(self, q, v):
        n = self.n
        q = np.asarray(q, np.float64)
        v = np.asarray(v, np.float64)
        pos = np.zeros((n, 3))
        rot = np.zeros((n, 4))
        joint = np.zeros((n, 3))
        axis = np.zeros((n, 3))
        if self.floating:
            pos[0] = q[:3]
            rot[0] = q[3:7] / np.linalg.norm(q[3:7])
        else:
            rot[0] = [0, 0, 0, 1]
        for i in range(1, n):
            b = self.bodies[i]
            p = int(b["parent"])
            pjr = qmul(rot[p], np.asarray(b["parent_rotation"], np.float64))
            an = np.asarray(b["axis"][:3], np.float64)
            an = an / np.linalg.norm(an)
            motion = np.array([0, 0, 0, 1.0])
            if int(b["q_index"]) != 0xFFFFFFFF:
                motion = axis_angle(an, q[int(b["q_index"])])
            r = qmul(qmul(pjr, motion), qconj(np.asarray(b["child_rotation"], np.float64)))
            rot[i] = r / np.linalg.norm(r)
            axis[i] = qrot(pjr, an)
            joint[i] = pos[p] + qrot(rot[p], np.asarray(b["parent_anchor"][:3], np.float64))
            pos[i] = joint[i] - qrot(rot[i], np.asarray(b["child_anchor"][:3], np.float64))
        return pos, rot, joint, axis

    def ancestors(self, i):
        out = []
        while i != 0:
            out.append(i)
            i = int(self.bodies[i]["parent"])
        return out

    def point_jacobian(self, link, x, pos, joint, axis):
        """3 x nv linear and 3 x nv angular Jacobian of world point x on link."""
        Jv = np.zeros((3, self.nv))
        Jw = np.zeros((3, self.nv))
        if self.floating:
            Jv[:, 0:3] = np.eye(3)
            for k in range(3):
                e = np.eye(3)[k]
                Jw[:, 3 + k] = e
                Jv[:, 3 + k] = np.cross(e, x - pos[0])
        for j in self.ancestors(link):
            vi = int(self.bodies[j]["v_index"])
            if vi == 0xFFFFFFFF:
                continue
            Jw[:, vi] = axis[j]
            Jv[:, vi] = np.cross(axis[j], x - joint[j])
        return Jv, Jw

    def world_inertia(self, i, rot):
        R = qmat(rot[i])
        Ib = np.asarray(self.bodies[i]["inertia"][:, :3], np.float64)
        return R @ Ib @ R.T

    def mass_matrix(self, q, armature):
        pos, rot, joint, axis = self.kinematics(q, np.zeros(self.nv))
        M = np.zeros((self.nv, self.nv))
        for i in range(self.n):
            Jv, Jw = self.point_jacobian(i, pos[i], pos, joint, axis)
            m = float(self.bodies[i]["mass"][0])
            M += m * Jv.T @ Jv + Jw.T @ self.world_inertia(i, rot) @ Jw
        M += np.diag(armature)
        return M

    def bias(self, q, v, damping=True):
        pos, rot, joint, axis = self.kinematics(q, v)
        n = self.n
        v = np.asarray(v, np.float64)
        w = np.zeros((n, 3))
        vl = np.zeros((n, 3))
        al = np.zeros((n, 3))
        ac = np.zeros((n, 3))
        if self.floating:
            vl[0] = v[0:3]
            w[0] = v[3:6]
        for i in range(1, n):
            b = self.bodies[i]
            p = int(b["parent"])
            qd = v[int(b["v_index"])] if int(b["v_index"]) != 0xFFFFFFFF else 0.0
            w[i] = w[p] + axis[i] * qd
            vo = vl[p] + np.cross(w[p], joint[i] - pos[p])
            vl[i] = vo + np.cross(w[i], pos[i] - joint[i])
            al[i] = al[p] + np.cross(w[p], axis[i] * qd)
            ao = ac[p] + np.cross(al[p], joint[i] - pos[p]) + np.cross(w[p], np.cross(w[p], joint[i] - pos[p]))
            ac[i] = ao + np.cross(al[i], pos[i] - joint[i]) + np.cross(w[i], np.cross(w[i], pos[i] - joint[i]))
        h = np.zeros(self.nv)
        for i in range(n):
            m = float(self.bodies[i]["mass"][0])
            Iw = self.world_inertia(i, rot)
            F = m * (ac[i] - self.gravity)
            T = Iw @ al[i] + np.cross(w[i], Iw @ w[i])
            if damping:
                F = F + float(self.bodies[i]["damping"][0]) * vl[i]
                T = T + float(self.bodies[i]["damping"][1]) * w[i]
            Jv, Jw = self.point_jacobian(i, pos[i], pos, joint, axis)
            h += Jv.T @ F + Jw.T @ T
        return h

    def forward_dynamics(self, q, v, tau, armature, damping=True):
        M = self.mass_matrix(q, armature)
        return np.linalg.solve(M, np.asarray(tau, np.float64) - self.bias(q, v, damping))

    def point_response(self, q, armature, link, x, f):
        pos, rot, joint, axis = self.kinematics(q, np.zeros(self.nv))
        Jv, _ = self.point_jacobian(link, np.asarray(x, np.float64), pos, joint, axis)
        return np.linalg.solve(self.mass_matrix(q, armature), Jv.T @ np.asarray(f, np.float64))

    def kinetic_energy(self, q, v, armature):
        v = np.asarray(v, np.float64)
        return 0.5 * v @ self.mass_matrix(q, armature) @ v

    def potential_energy(self, q):
        pos, _, _, _ = self.kinematics(q, np.zeros(self.nv))
        return -sum(float(self.bodies[i]["mass"][0]) * self.gravity @ pos[i] for i in range(self.n))
