"""Mixed-size persistent solver (fadesim_ragged_*) on the GPU.

* fixed sweeps, mixed sizes including 256x256 and 512x512 (strips with row
  exchange between workgroups): every instance's drop field is bitwise equal
  to the CPU FP32 red-black replay (oracle_rb_f32), from zero and random warm
  starts -- this pins the strip exchange;
* converged, the uneven-size workload's distribution: every instance within
  1e-6 relative bitline current of the exact FP64 solve (oracle.line);
* a batch of only 128x128 crossbars: bitwise the results of the uniform
  solver (same method, schedule and relaxation);
* the launch captured into a graph and replayed, with eager work between
  replays, reproduces the eager results exactly;
* the in-kernel IR2S guess on strips (256x256, 512x512: column sums over the
  team) against the FP64 drop-form oracle, and damped IR2S guesses with
  converged solves against the exact solve.
"""
import numpy as np
import pytest

import oracle
from oracle.line import line_solve

pytestmark = pytest.mark.gpu


def _mixed(cuda, sizes, seed, numpy_too=True):
    import torch
    from fadesim_amd.workloads import uneven_batch
    counts = {}
    for n in sizes:
        counts[n] = counts.get(n, 0) + 1
    return uneven_batch(counts, seed=seed, device=cuda, numpy_too=numpy_too)


def test_fixed_sweeps_bitwise_vs_replay(cuda):
    import torch
    from fadesim_amd.ragged import RaggedPlan, SIZES
    inst = _mixed(cuda, [64, 128, 256, 512, 256, 64, 512, 128, 256], seed=3)
    plan = RaggedPlan(inst["rows"], device=cuda, tuning={})  # plain SOR: no IR2S guess
    rng = np.random.default_rng(9)
    warm_list = []
    for k, n in enumerate(inst["rows"]):
        w = np.zeros((2, n, n), np.float32)
        if k % 2:
            w[:] = rng.uniform(0, 1e-3, (2, n, n)).astype(np.float32)
        warm_list.append(w)
    warm = torch.tensor(plan.pack(warm_list, "field", np.float32), device=cuda)
    sweeps = 37
    res = plan.solve(inst["g"], inst["r"], inst["vin"], warm=warm, fixed_sweeps=sweeps, want_field=True)
    torch.cuda.synchronize()
    fields = plan.unpack(res.field, "field")
    assert (res.sweeps.cpu().numpy() == sweeps).all()
    for k, n in enumerate(inst["rows"]):
        w = plan.stats["omega"][SIZES.index(n)]
        p, q = oracle.rb_f32(inst["g_list"][k], inst["r_np"][k], inst["vin_list"][k], None, w, sweeps,
                             warm_list[k][0], warm_list[k][1])
        assert np.array_equal(fields[k][0], p), (k, n)
        assert np.array_equal(fields[k][1], q), (k, n)


def test_converged_mixed_vs_exact(cuda):
    from fadesim_amd.ragged import RaggedPlan
    inst = _mixed(cuda, [64, 64, 128, 128, 256, 256, 512, 512, 128, 64], seed=11)
    plan = RaggedPlan(inst["rows"], device=cuda)
    res = plan.solve(inst["g"], inst["r"], inst["vin"], tol=1e-6)
    st = res.status.cpu().numpy()
    assert (st == 0).all(), st
    iouts = plan.unpack(res.iout, "cols")
    for k, n in enumerate(inst["rows"]):
        _, _, iex, _ = line_solve(inst["g_list"][k], inst["r_np"][k], inst["vin_list"][k])
        err = np.max(np.abs(iouts[k] - iex) / np.abs(iex))
        assert err <= 1e-6, (k, n, err)


def test_uniform_128_batch_matches_uniform_solver(cuda):
    import torch
    from fadesim_amd.ragged import RaggedPlan
    from fadesim_amd.solver import solve_batch
    from fadesim_amd.workloads import uniform_batch
    g, r, vin = uniform_batch(40, 128, seed=5, device=cuda)
    plan = RaggedPlan([128] * 40, device=cuda, tuning={})
    res = plan.solve(g.reshape(-1).contiguous(), r, vin.reshape(-1).contiguous(), tol=1e-6)
    ref = solve_batch(g, r, vin, tol=1e-6)
    assert torch.equal(res.iout.view(40, 128), ref.iout)
    assert torch.equal(res.sweeps, ref.sweeps)


def test_graph_replays_match_eager(cuda):
    import torch
    from fadesim_amd.ragged import RaggedPlan
    inst = _mixed(cuda, [64, 128, 256, 512, 128], seed=17, numpy_too=False)
    plan = RaggedPlan(inst["rows"], device=cuda)
    eager = plan.solve(inst["g"], inst["r"], inst["vin"], tol=1e-6)
    e_iout, e_sw = eager.iout.clone(), eager.sweeps.clone()
    graph, res = plan.capture(inst["g"], inst["r"], inst["vin"], tol=1e-6)
    for rep in range(3):
        res.iout.fill_(0.0)
        graph.replay()
        torch.cuda.synchronize()
        assert not plan.aborted()
        assert torch.equal(res.iout, e_iout), rep
        assert torch.equal(res.sweeps, e_sw), rep
        # eager work between replays
        plan.solve(inst["g"], inst["r"], inst["vin"], tol=1e-6)
        torch.cuda.synchronize()


def test_plan_rejects_unsupported_sizes(cuda):
    from fadesim_amd.ragged import RaggedPlan
    with pytest.raises(RuntimeError):
        RaggedPlan([64, 96], device=cuda)
    with pytest.raises(RuntimeError):
        RaggedPlan([128], cols=[64], device=cuda)


@pytest.mark.parametrize("iters,omega", [(3, 0.8), (5, -1.1)])
def test_ir2s_guess_on_strips_matches_oracle(cuda, iters, omega):
    """Damped and Chebyshev-scheduled (omega < 0: the bound from the
    conductance summed over the strips of a team) wire-solve guesses."""
    from fadesim_amd.ragged import RaggedPlan, SIZES
    from oracle.drops import ir2s_drops
    inst = _mixed(cuda, [256, 512, 64, 128], seed=21)
    tuning = {str(n): {"ir2s_iters": iters, "ir2s_omega": omega} for n in SIZES}
    plan = RaggedPlan(inst["rows"], device=cuda, tuning=tuning)
    res = plan.solve(inst["g"], inst["r"], inst["vin"], fixed_sweeps=0, want_field=True)
    fields = plan.unpack(res.field, "field")
    for k, n in enumerate(inst["rows"]):
        p, q, _ = ir2s_drops(inst["g_list"][k].astype(np.float64), inst["r_np"][k], inst["vin_list"][k],
                             None, iters, omega)
        for a, b in ((fields[k][0], p), (fields[k][1], q)):
            err = np.max(np.abs(a.astype(np.float64) - b)) / np.abs(b).max()
            assert err <= 2e-5, (k, n, err)


def test_ir2s_damped_converged_mixed_vs_exact(cuda):
    from fadesim_amd.ragged import RaggedPlan, SIZES
    inst = _mixed(cuda, [64, 128, 256, 512, 512, 256], seed=23)
    tuning = {"64": {"ir2s_iters": 4, "corr_iters": 4}, "128": {"ir2s_iters": 6, "corr_iters": 6},
              "256": {"ir2s_iters": 8, "corr_iters": 8, "ir2s_omega": 0.85},
              "512": {"ir2s_iters": 12, "corr_iters": 12, "ir2s_omega": 0.7}}
    plan = RaggedPlan(inst["rows"], device=cuda, tuning=tuning)
    res = plan.solve(inst["g"], inst["r"], inst["vin"], tol=1e-6)
    assert (res.status.cpu().numpy() == 0).all()
    base = RaggedPlan(inst["rows"], device=cuda, tuning={}).solve(inst["g"], inst["r"], inst["vin"], tol=1e-6)
    iouts = plan.unpack(res.iout, "cols")
    for k, n in enumerate(inst["rows"]):
        _, _, iex, _ = line_solve(inst["g_list"][k], inst["r_np"][k], inst["vin_list"][k])
        err = np.max(np.abs(iouts[k] - iex) / np.abs(iex))
        assert err <= 1e-6, (k, n, err)
    assert res.sweeps.double().sum().item() < 0.5 * base.sweeps.double().sum().item()



def test_shipped_tuning_table_converged_vs_exact(cuda):
    """The plan with the committed table for this GPU (the Chebyshev-scheduled
    guesses: 2 / 3 / 5 / 8 iterations by size on gfx950): every crossbar of a
    mixed batch within 1e-6 of the exact solve."""
    from fadesim_amd.ragged import RaggedPlan
    inst = _mixed(cuda, [512, 64, 256, 128, 512, 256, 128, 64], seed=29)
    plan = RaggedPlan(inst["rows"], device=cuda)
    res = plan.solve(inst["g"], inst["r"], inst["vin"], tol=1e-6)
    assert (res.status.cpu().numpy() == 0).all()
    iouts = plan.unpack(res.iout, "cols")
    for k, n in enumerate(inst["rows"]):
        _, _, iex, _ = line_solve(inst["g_list"][k], inst["r_np"][k], inst["vin_list"][k])
        err = np.max(np.abs(iouts[k] - iex) / np.abs(iex))
        assert err <= 1e-6, (k, n, err)

def test_strip_exchange_with_every_cu_occupied(cuda):
    """48 crossbars of 512x512 and 64 of 256x256 among 128x128 and 64x64 ones:
    more team tasks than the grid has teams, so every workgroup of the
    persistent grid is busy and the teams' exchanges run side by side with
    the small items on every CU.  Converged with the tuned schedule (IR2S
    guess, team column sums, refinement), each sampled crossbar's outputs
    equal bit for bit its solve alone in a plan of its own; with fixed sweeps
    the sampled strips' drop fields equal the CPU FP32 red-black replay."""
    import torch
    from fadesim_amd.ragged import RaggedPlan, SIZES
    inst = _mixed(cuda, [512] * 48 + [256] * 64 + [128] * 96 + [64] * 96, seed=41)
    plan = RaggedPlan(inst["rows"], device=cuda)
    assert plan.stats["big_tasks"] > plan.stats["teams"]
    res = plan.solve(inst["g"], inst["r"], inst["vin"], tol=1e-6)
    torch.cuda.synchronize()
    assert not plan.aborted()
    assert (res.status.cpu().numpy() == 0).all()
    iouts = plan.unpack(res.iout, "cols")
    rows = list(inst["rows"])
    picks = [rows.index(512), len(rows) - 1 - rows[::-1].index(512), rows.index(256),
             len(rows) - 1 - rows[::-1].index(256), rows.index(128), rows.index(64)]
    for k in picks:
        n = rows[k]
        one = RaggedPlan([n], device=cuda)
        g1 = torch.tensor(inst["g_list"][k].reshape(-1), device=cuda)
        r1 = torch.tensor([inst["r_np"][k]], dtype=torch.float64, device=cuda)
        v1 = torch.tensor(inst["vin_list"][k], dtype=torch.float64, device=cuda)
        alone = one.solve(g1, r1, v1, tol=1e-6)
        torch.cuda.synchronize()
        assert np.array_equal(alone.iout.cpu().numpy(), iouts[k]), (k, n)
        assert int(alone.sweeps[0]) == int(res.sweeps[k]), (k, n)
    # fixed sweeps: the row exchange of every strip under full occupancy
    plain = RaggedPlan(inst["rows"], device=cuda, tuning={})
    sweeps = 23
    fx = plain.solve(inst["g"], inst["r"], inst["vin"], fixed_sweeps=sweeps, want_field=True)
    torch.cuda.synchronize()
    assert not plain.aborted()
    fields = plain.unpack(fx.field, "field")
    for k in picks[:4]:
        n = rows[k]
        w = plain.stats["omega"][SIZES.index(n)]
        p, q = oracle.rb_f32(inst["g_list"][k], inst["r_np"][k], inst["vin_list"][k], None, w, sweeps)
        assert np.array_equal(fields[k][0], p), (k, n)
        assert np.array_equal(fields[k][1], q), (k, n)
