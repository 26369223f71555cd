"""Step server (metalrobo.server): requests served from a resident pool of
worlds match the same worlds stepped on their own, bit for bit; admission
waits for free slots and rejects requests larger than the pool; the Poisson
request generator follows its distribution."""

import numpy as np
import pytest

import metalrobo as mr
from metalrobo import server as sv


def test_poisson_requests_follow_their_distribution():
    reqs = sv.poisson_requests(200.0, 4000, seed=1)
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    assert abs(gaps.mean() - 1.0 / 200.0) < 0.1 / 200.0
    n = np.array([r.worlds for r in reqs])
    h = np.array([r.horizon for r in reqs])
    assert n.min() >= 1 and n.max() <= 1024 and h.min() >= 10 and h.max() <= 500
    # log-uniform: the median sits near sqrt(1 * 1024) = 32
    assert 20 <= np.median(n) <= 50


@pytest.mark.gpu
def test_served_worlds_match_the_same_worlds_stepped_alone(gpu):
    pool = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, 48, 3)
    srv = sv.StepServer(pool, hbm_budget_bytes=200e9, command_bank=8)
    reqs = [sv.Request(0, 5, 4, 0.0), sv.Request(1, 9, 6, 0.0), sv.Request(2, 3, 3, 0.0),
            sv.Request(3, 40, 2, 0.0)]  # the last one waits for slots
    rep = srv.serve(reqs, return_states=True)
    assert rep.env_steps == sum(r.worlds * r.horizon for r in reqs)
    assert reqs[3].admitted >= reqs[2].finished
    assert set(rep.results) == {0, 1, 2, 3}
    cmds = srv.commands.cpu().numpy()
    # requests 0-2 start at tick 0, so their control step t used bank row t
    for r in reqs[:3]:
        alone = mr.World(pool, pool.default_config())
        for t in range(r.horizon):
            alone.advance(cmds[t % cmds.shape[0]:t % cmds.shape[0] + 1])
        qa, va, sa = alone.state()
        slots = r.slots.astype(np.int64)
        q, v, s = rep.results[r.rid]
        assert np.array_equal(q, qa[slots]) and np.array_equal(v, va[slots])
        assert np.array_equal(s.view(np.uint8), sa[slots].view(np.uint8))
        alone.close()


@pytest.mark.gpu
def test_oversized_requests_are_rejected(gpu):
    pool = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, 8, 3)
    srv = sv.StepServer(pool, command_bank=2)
    with pytest.raises(ValueError, match="more worlds"):
        srv.serve([sv.Request(0, 9, 2, 0.0)])


@pytest.mark.gpu
def test_served_poisson_load_at_scale_binds_the_budget_and_matches_the_oracle(gpu):
    """The step-serving configuration at its own scale: Poisson arrivals at 200 requests/s of
    N log-uniform in [1, 1024] humanoid worlds on a 4096-slot pool, admitted under an HBM budget
    of 1536 slots' bytes (below the pool's), so admission waits on the budget. Horizons are
    drawn from [10, 40] control steps to keep the oracle check short. Sampled worlds of served
    requests are stepped by the oracle from the reset state with the command-bank rows their
    ticks used, and must end bit for bit where the server left them."""
    import oracle_lib as ol

    E = 4096
    pool = mr.WorldModel.synthetic(mr.MR_SYNTHETIC_HUMANOID_TERRAIN, E, 3)
    cfg = pool.default_config()
    srv = sv.StepServer(pool, cfg, command_bank=16)
    srv.set_hbm_budget(1536 * srv.bytes_per_slot)
    assert srv.capacity == 1536 and srv.hbm_budget < E * srv.bytes_per_slot
    reqs = sv.poisson_requests(200.0, 32, seed=5, horizon=(10, 40))
    print(f"serving {len(reqs)} requests, {sum(r.worlds for r in reqs)} worlds", flush=True)
    rep = srv.serve(reqs, return_states=True)
    assert rep.env_steps == sum(r.worlds * r.horizon for r in reqs)
    assert set(rep.results) == {r.rid for r in reqs}
    assert rep.budget_blocked_ticks > 0, "the budget never held a request back"
    assert 0 < rep.max_admitted_bytes <= srv.hbm_budget
    assert rep.max_active_worlds <= 1536
    assert rep.latency_ms_p99 >= rep.latency_ms_p50 > 0.0
    cmds = srv.commands.cpu().numpy()
    bank = cmds.shape[0]
    st = ol.OracleState(pool, cfg)
    checked = 0
    for r in reqs[::5]:
        for s in {int(r.slots[0]), int(r.slots[-1])}:
            for t in range(r.horizon):
                mask = np.zeros(E, np.uint32)
                mask[s] = 1 if t == 0 else 0
                status, *_ = st.step(cmds[(r.first_tick + t) % bank], reset_mask=mask, env_begin=s, env_end=s + 1)
                assert status["code"][s] == 0
            q, v, scene = rep.results[r.rid]
            row = int(np.nonzero(r.slots == s)[0][0])
            assert np.array_equal(q[row].view(np.uint32), st.q[s].view(np.uint32)), (r.rid, s)
            assert np.array_equal(v[row].view(np.uint32), st.v[s].view(np.uint32)), (r.rid, s)
            checked += 1
    assert checked >= 8
