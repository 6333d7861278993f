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


def pooled_trace(zh, torch, count, rate_per_s, seed, sides, own_every):
    """A trace whose requests write to the server's output pool (out = 0),
    except every `own_every`-th one, which gets a buffer of its own to check."""
    chosen, arrivals, ins, outs, specs = trace(zh, torch, count, rate_per_s, seed, sides)
    checked = list(range(0, count, own_every))
    specs = [(di, do if i in checked else 0, k, c, n) for i, (di, do, k, c, n) in enumerate(specs)]
    return chosen, arrivals, ins, outs, specs, checked


def test_pool_bounded_by_budget(zh, torch):
    sides = (64, 256, 512)
    chosen, arrivals, ins, outs, specs, checked = pooled_trace(zh, torch, 30, 20000, 31, sides, 3)
    out_max = max(2 * s * 128 * s * 4 for s in chosen)
    budget = 3 * out_max                       # room for a few outputs at a time
    print(f"[case] pooled trace budget={budget} sides={chosen}", flush=True)
    lat, stats = zh.serve_trace(specs, arrivals, hbm_budget=budget, slots=3, mode=0)
    print(f"[serve] pooled stats={stats}", flush=True)
    assert (lat > 0).all() and stats["rejected"] == 0
    assert 0 < stats["pool_peak_bytes"] <= budget
    assert stats["peak_bytes"] <= budget
    for i in checked:
        assert torch.equal(outs[i], alone(zh, torch, ins[i])), (i, chosen[i])
    # the plain loop takes the same trace (one output block per size)
    lat1, stats1 = zh.serve_trace(specs, arrivals, hbm_budget=budget, slots=3, mode=1)
    assert (lat1 > 0).all()
    for i in checked:
        assert torch.equal(outs[i], alone(zh, torch, ins[i])), (i, chosen[i])


def test_online_pool_handoff(zh, torch):
    import time
    sides, _, ins, _, specs = trace(zh, torch, 12, 1000, seed=41, sides=(64, 128, 256))
    out_max = max(2 * s * 128 * s * 4 for s in sides)
    in_max = max(s * 128 * s * 4 for s in sides)
    budget = 2 * out_max + in_max
    srv = zh.Server(hbm_budget=budget, slots=2)
    tickets = [srv.submit(di, 0, k, c, n) for (di, _, k, c, n) in specs]
    pending = dict(zip(tickets, range(len(tickets))))
    deadline = time.time() + 60
    while pending and time.time() < deadline:
        for t in list(pending):
            state, lat = srv.poll(t)
            if state != 2:
                continue
            i = pending.pop(t)
            s = sides[i]
            got = zh.device_view(srv.output(t), 2 * s, 128 * s).clone()
            torch.cuda.synchronize()  # the copy has read the block before it goes back
            srv.release(t)
            assert lat > 0
            assert torch.equal(got, alone(zh, torch, ins[i])), (i, s)
    assert not pending, f"{len(pending)} requests not done"
    stats = srv.stats()
    assert 0 < stats["pool_peak_bytes"] <= budget
    # a finished, polled and released request is forgotten
    with pytest.raises(zh.ZodaError):
        srv.poll(tickets[0])
    srv.close()


def test_unreleased_outputs_are_reported(zh, torch):
    # Outputs held by finished requests that are never released block the
    # queue once the pool is full: wait_all says so instead of waiting forever.
    sides, _, ins, _, specs = trace(zh, torch, 4, 1000, seed=43, sides=(256,))
    out_b = 2 * 256 * 128 * 256 * 4
    srv = zh.Server(hbm_budget=out_b + 256 * 128 * 256 * 4, slots=1)
    for (di, _, k, c, n) in specs:
        srv.submit(di, 0, k, c, n)
    with pytest.raises(zh.ZodaError) as e:
        srv.wait_all()
    assert e.value.code == -4
    srv.close()
