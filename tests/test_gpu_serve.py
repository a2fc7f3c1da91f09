"""The request loop of the served form: requests admitted while others run
come out bit-identical to the same scene stepped alone."""

import numpy as np
import pytest

from emerald_amd import serve
from emerald_amd import sph2d_box as sb
from helpers import assert_bits_equal

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone(gpu):
    # Small scenes, short requests and a tight budget, so that scenes are
    # admitted into running pools, retired, and wait for room.
    reqs = serve.poisson_requests(400.0, 1.5, seed=9, n_min=300, n_max=3000, steps_min=3,
                                  steps_max=40)
    print(f"[serve parity] {len(reqs)} requests", flush=True)
    res = serve.serve(reqs, budget_gb_to_bytes(0.15), 2.0, keep_states=True)
    try:
        loops = res["_loops"]
        done = {rid: st for lp in loops.values() for rid, st in lp.finals.items()}
        assert res["completed"] >= 10 and len(done) == res["completed"]
        assert res["admitted"] > res["completed"] or res["live_at_end"] == 0
        by_id = {r.rid: r for r in reqs}
        checked = 0
        for rid in sorted(done)[:12]:
            req = by_id[rid]
            sc, cfg = serve.request_scene(req)
            ctx = sb.Context(cfg, sb.Solid_state(sc.solids), len(sc.positions))
            ctx.set_cfl(0.4)
            ctx.set_state(sb.State(sc.positions, sc.velocities))
            ctx.frame_init(req.method)
            ctx.run_sub_steps(req.method, cfg.seconds_per_sub_step, req.steps)
            want = ctx.get_state()
            ctx.close()
            assert_bits_equal(done[rid].positions, want.positions, f"request {rid} positions")
            assert_bits_equal(done[rid].velocities, want.velocities, f"request {rid} velocities")
            checked += 1
        assert checked >= 10
    finally:
        serve.close(res)


def budget_gb_to_bytes(gb):
    return gb * 1e9
