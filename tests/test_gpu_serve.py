"""Serving loop: every result of the continuous-batching server equals the
same square encoded alone, for the trace replay and the online API; the HBM
budget bounds what is admitted at once and rejects what can never fit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def zh():
    import zoda_hip
    assert zoda_hip.gpu_available(), "no HIP device"
    return zoda_hip


@pytest.fixture(scope="module")
def torch():
    import torch
    return torch


def trace(zh, torch, count, rate_per_s, seed, sides=(64, 128, 256, 512, 1024)):
    rng = np.random.default_rng(seed)
    chosen = rng.choice(sides, size=count).tolist()            # log-uniform over the sides
    gaps = rng.exponential(1e6 / rate_per_s, size=count)       # Poisson arrivals
    arrivals = np.cumsum(gaps) - gaps[0]
    ins, outs, specs = [], [], []
    for i, s in enumerate(chosen):
        cols = 128 * s
        t_in = torch.empty((s, cols), dtype=torch.int32, device="cuda")
        zh.fill_uniform(t_in, seed=seed * 10000 + i)
        t_out = torch.zeros((2 * s, cols), dtype=torch.int32, device="cuda")
        ins.append(t_in)
        outs.append(t_out)
        specs.append((t_in.data_ptr(), t_out.data_ptr(), s, cols, s))
    torch.cuda.synchronize()
    return chosen, arrivals, ins, outs, specs


def alone(zh, torch, t_in):
    out = zh.encode_tensor(t_in)
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("mode", [0, 1])
def test_trace_results_equal_alone(zh, torch, mode):
    sides, arrivals, ins, outs, specs = trace(zh, torch, 40, 2000, seed=5 + mode)
    print(f"[case] serve mode={mode} sides={sides}", flush=True)
    lat, stats = zh.serve_trace(specs, arrivals, hbm_budget=200 * 10**9, slots=3, mode=mode)
    print(f"[serve] mode={mode} stats={stats} p50={np.percentile(lat, 50):.1f}us "
          f"p99={np.percentile(lat, 99):.1f}us", flush=True)
    assert (lat > 0).all()
    assert stats["rejected"] == 0
    for i, (t_in, t_out) in enumerate(zip(ins, outs)):
        assert torch.equal(t_out, alone(zh, torch, t_in)), (i, sides[i])


def test_burst_of_new_sides(zh, torch):
    # Requests of sides the server has not seen yet, a few microseconds
    # apart: consecutive batched launches land on different slot streams, and
    # each must see the table entry its square needs (uploaded before any
    # launch names it), whichever stream launched first.
    rng = np.random.default_rng(21)
    sides = [int(s) for s in rng.choice([16, 32, 64, 128], size=24)]
    ins, outs, specs = [], [], []
    for i, s in enumerate(sides):
        t_in = torch.empty((s, 128 * s), dtype=torch.int32, device="cuda")
        zh.fill_uniform(t_in, seed=700 + i)
        t_out = torch.zeros((2 * s, 128 * s), dtype=torch.int32, device="cuda")
        ins.append(t_in)
        outs.append(t_out)
        specs.append((t_in.data_ptr(), t_out.data_ptr(), s, 128 * s, s))
    torch.cuda.synchronize()
    arrivals = np.arange(len(sides), dtype=np.float64) * 5.0
    print(f"[case] serve burst sides={sides}", flush=True)
    lat, stats = zh.serve_trace(specs, arrivals, hbm_budget=200 * 10**9, slots=4, mode=0)
    assert (lat > 0).all() and stats["rejected"] == 0
    for i, (t_in, t_out) in enumerate(zip(ins, outs)):
        assert torch.equal(t_out, alone(zh, torch, t_in)), (i, sides[i])


def test_budget_bounds_admission(zh, torch):
    sides, arrivals, ins, outs, specs = trace(zh, torch, 24, 20000, seed=9, sides=(64, 256, 512))
    biggest = max((s + 2 * s) * 128 * s * 4 for s in sides)
    lat, stats = zh.serve_trace(specs, arrivals, hbm_budget=biggest, slots=4, mode=0)
    assert stats["peak_bytes"] <= biggest
    assert (lat > 0).all()
    for t_in, t_out in zip(ins, outs):
        assert torch.equal(t_out, alone(zh, torch, t_in))


def test_online_server(zh, torch):
    sides, _, ins, outs, specs = trace(zh, torch, 20, 1000, seed=11, sides=(64, 128, 256))
    srv = zh.Server(hbm_budget=64 * 10**9, slots=2)
    tickets = [srv.submit(*spec) for spec in specs]
    srv.wait_all()
    for t, t_in, t_out in zip(tickets, ins, outs):
        state, lat = srv.poll(t)
        assert state == 2 and lat > 0
        assert torch.equal(t_out, alone(zh, torch, t_in))
    stats = srv.stats()
    assert stats["batches"] >= 1 and stats["rejected"] == 0
    # a request that can never fit the budget is rejected, not queued forever
    small = zh.Server(hbm_budget=1 << 20, slots=1)
    with pytest.raises(zh.ZodaError) as e:
        small.submit(*specs[-1])
    assert e.value.code == -4
    small.close()
    srv.close()
