"""Pins the oracle's relaxed block Jacobi cone solve (the Wave32 production
iteration, mrWorldWave32SolvePacket, MetalWorldContact.metal:6142-7213) by
cases whose answers do not come from the iteration's own formulas (CPU only):

  * one contact: sum_j |W_nn(c, j)| is the contact's own term, so the damping
    is exactly 1 and applying "all deltas at once" is applying the one delta;
    the trajectory must equal the ordered sweeps' bit for bit;
  * two stacked spheres (1 kg on the plane, 3 kg falling onto it), lever arms
    along the normals and no regularisation: the sphere-sphere contact couples
    to the sphere-plane contact only through the 1 kg sphere, so its damping
    is (1/1 + 1/3) / (1/1 + 1/3 + 1/1) = 4/7, and its first Jacobi impulse
    is 4/7 of the first ordered-sweep impulse from the same state;
  * converged from either iteration, a box resting on a plane carries its
    weight (the fixed point of both iterations is the contact solution).
"""

import numpy as np
import pytest

import metalrobo as mr
import model_util as mu
import oracle_lib as ol
from test_oracle_solver import box_on_plane, run, sliding_sphere

JACOBI = mr.MR_SOLVER_RELAXED_JACOBI
SWEEPS = mr.MR_SOLVER_ORDERED_SWEEPS
PLANE = 3


def test_single_contact_jacobi_equals_ordered_sweeps():
    states = []
    for solver in (SWEEPS, JACOBI):
        w, cfg, st = sliding_sphere(1.0, 0.7, (1.5, 0.5, -0.2))
        cfg.solver = solver
        run(st, 8)
        states.append(st.scene.copy())
    np.testing.assert_array_equal(states[0].view(np.uint8), states[1].view(np.uint8))


def stacked_spheres(solver):
    bodies = [mu.body(1.0, [0.01, 0.01, 0.01]), mu.body(1.0, [0.004] * 3), mu.body(3.0, [0.012] * 3),
              mu.body(0.0, [1, 1, 1], motion=0)]
    shapes = [mu.shape(1, 0, [0.1]), mu.shape(2, 0, [0.1]), mu.shape(3, PLANE, [0, 0, 0], rot=mu.Y_TO_Z)]
    mats = [((0.8, 0.8, 0.8, 0.8), (0.0, 0.5, 0.0, 0.0))]
    w = mu.CustomWorld(bodies, [], shapes, link_count=1, root_type=0, nq=0, nv=0, materials=mats)
    cfg = w.config(physics_substeps=1, control_timestep=0.005, velocity_iterations=1, final_velocity_iterations=0,
                   warm_start=0)
    cfg.solver = solver
    st = ol.OracleState(w.model, cfg)
    st.scene[0, 0] = mu.scene_state((0, 0, 0.1), inv_mass=1.0)
    st.scene[0, 1] = mu.scene_state((0, 0, 0.3), lin=(0, 0, -1.0), inv_mass=1.0 / 3.0)
    st.scene[0, 2] = mu.scene_state((0, 0, 0), motion=0)
    return w, cfg, st


def test_two_sphere_stack_damping_is_the_coupling_ratio():
    first = {}
    for solver in (SWEEPS, JACOBI):
        w, cfg, st = stacked_spheres(solver)
        run(st, 1)
        assert int(st.mcount[0]) == 2
        # the sphere-sphere manifold comes first (compiled pair order) and holds one point
        assert tuple(st.mh[0, 0]["pair_and_count"][1:4]) == (0, 1, 1)
        assert tuple(st.mh[0, 1]["pair_and_count"][1:3]) == (0, 2)
        first[solver] = float(st.mp[0, 0, 0]["impulses"][0])
    assert first[SWEEPS] > 0.0
    ratio = first[JACOBI] / first[SWEEPS]
    assert abs(ratio - 4.0 / 7.0) < 1.0e-5, ratio


@pytest.mark.parametrize("solver", [SWEEPS, JACOBI])
def test_converged_resting_box_carries_its_weight(solver):
    w, cfg, st = box_on_plane()
    cfg.solver = solver
    cfg.residual_tolerance = 1.0e-9
    cfg.max_iterations = 512
    run(st, 50)
    h = cfg.control_timestep / cfg.physics_substeps
    assert int(st.mcount[0]) == 1
    lam = st.mp[0, 0]["impulses"][:4, 0].sum()
    assert abs(lam - 9.81 * h) / (9.81 * h) < 0.02, lam
    assert np.linalg.norm(st.scene[0, 0]["linear_velocity_inv_mass"][:3]) < 1e-3
