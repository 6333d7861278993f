"""GPU parity of the batched persistent encode and its hipGraph replay: every
square of a mixed batch must equal its plain single-square encode (and the
oracle), alone and in the mix, for every autotuner variant."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
P = O.P


@pytest.fixture(scope="module")
def zh():
    import zoda_hip
    assert zoda_hip.gpu_available(), "no HIP device"
    return zoda_hip


@pytest.fixture(scope="module")
def torch():
    import torch
    return torch


def mixed_sides(count, seed):
    # the batched workload: sides from {32..512} weighted 30/30/20/15/5 %
    rng = np.random.default_rng(seed)
    return rng.choice([32, 64, 128, 256, 512], size=count, p=[0.30, 0.30, 0.20, 0.15, 0.05]).tolist()


def make_batch(zh, torch, sides, seed, col_scale=128, ragged=False):
    ins, outs, specs = [], [], []
    for i, s in enumerate(sides):
        cols = s * col_scale + (7 * i + 3 if ragged else 0)
        k = s - (i % 3 if ragged and s > 4 else 0)
        t_in = torch.empty((k, cols), dtype=torch.int32, device="cuda")
        zh.fill_uniform(t_in, seed=seed * 1000 + i)
        t_out = torch.empty((2 * s, cols), dtype=torch.int32, device="cuda")
        ins.append(t_in)
        outs.append(t_out)
        specs.append((t_in.data_ptr(), t_out.data_ptr(), k, cols, s))
    return ins, outs, specs


def plain(zh, torch, t_in, n):
    out = zh.encode_tensor(t_in, n=n)
    torch.cuda.synchronize()
    return out


def test_batch_equals_plain_and_oracle(zh, torch):
    sides = mixed_sides(96, 1)
    print(f"[case] batch96 sides={sides}", flush=True)
    ins, outs, specs = make_batch(zh, torch, sides, seed=1)
    zh.encode_batch(specs, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    for i, (t_in, t_out, s) in enumerate(zip(ins, outs, sides)):
        assert torch.equal(t_out, plain(zh, torch, t_in, s)), (i, s)
    # oracle on a few whole squares
    for i in [sides.index(x) for x in sorted(set(sides)) if x <= 128]:
        data = ins[i].cpu().numpy().view(np.uint32)
        assert np.array_equal(outs[i].cpu().numpy().view(np.uint32), O.encode(data, threads=8)), sides[i]


def test_batch_ragged_shapes(zh, torch):
    sides = [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096]
    ins, outs, specs = make_batch(zh, torch, sides, seed=2, col_scale=2, ragged=True)
    zh.encode_batch(specs)
    torch.cuda.synchronize()
    for t_in, t_out, s in zip(ins, outs, sides):
        data = t_in.cpu().numpy().view(np.uint32)
        expect = O.encode(data, n=s, threads=8) if data.shape[0] > 0 else None
        assert np.array_equal(t_out.cpu().numpy().view(np.uint32), expect), s


def test_graph_replay(zh, torch):
    sides = mixed_sides(24, 3)
    ins, outs, specs = make_batch(zh, torch, sides, seed=3)
    plan = zh.BatchPlan(specs)
    stream = torch.cuda.current_stream().cuda_stream
    for rep in range(4):
        for i, t_in in enumerate(ins):  # new contents, eager work between replays
            zh.fill_uniform(t_in, seed=100 * rep + i)
        torch.cuda.synchronize()
        for t_out in outs:
            t_out.zero_()
        plan.launch(stream)
        torch.cuda.synchronize()
        for i, (t_in, t_out, s) in enumerate(zip(ins, outs, sides)):
            assert torch.equal(t_out, plain(zh, torch, t_in, s)), (rep, i, s)
    plan.destroy()


@pytest.mark.parametrize("sides", [[512, 256, 256], [32, 64, 128]])
def test_graph_replay_one_kind(zh, torch, sides):
    # only squares that run as their own fused launches, or only queued ones
    ins, outs, specs = make_batch(zh, torch, sides, seed=4)
    plan = zh.BatchPlan(specs)
    stream = torch.cuda.current_stream().cuda_stream
    for rep in range(2):
        for t_out in outs:
            t_out.zero_()
        plan.launch(stream)
        torch.cuda.synchronize()
        for i, (t_in, t_out, s) in enumerate(zip(ins, outs, sides)):
            assert torch.equal(t_out, plain(zh, torch, t_in, s)), (rep, i, s)
    plan.destroy()


def test_autotune_variants_agree(zh, torch, tmp_path):
    tuned = zh.autotune([32, 64, 128, 256, 512, 1024], iters=2)
    print("[autotune]", {s: (e["choice"], e["times_us"]) for s, e in tuned.items()}, flush=True)
    for s, e in tuned.items():
        assert len(e["times_us"]) == zh.TUNE_CANDIDATES and all(0 < t < 1e29 for t in e["times_us"]), (s, e)
        best = min(range(zh.TUNE_CANDIDATES), key=lambda v: e["times_us"][v])
        assert e["times_us"][e["choice"]] == e["times_us"][best], (s, e)
    # the table survives a save and load
    path = str(tmp_path / "tune.txt")
    zh.tune_save(path)
    for s in tuned:
        zh.set_variant(s, 0)
    assert zh.tune_load(path) == 13
    for s, e in tuned.items():
        assert zh.get_variant(s) == e["choice"], s
    # every candidate (four queue plans, own launches) gives the same squares
    sides = [16, 32, 64, 128, 256, 512, 1024]
    ins, outs, specs = make_batch(zh, torch, sides, seed=4, col_scale=4)
    results = []
    for v in range(zh.TUNE_CANDIDATES):
        print(f"[case] batch candidate {v}", flush=True)
        for s in sides:
            zh.set_variant(s, v)
        for o in outs:
            o.zero_()
        zh.encode_batch(specs)
        torch.cuda.synchronize()
        results.append([o.clone() for o in outs])
    for i, (s, t_in) in enumerate(zip(sides, ins)):
        ref = plain(zh, torch, t_in, s)
        for v in range(zh.TUNE_CANDIDATES):
            assert torch.equal(results[v][i], ref), (s, v)
    for s, entry in tuned.items():
        zh.set_variant(s, entry["choice"])


def test_autotune_for_a_batch(zh, torch):
    sides = [32, 64, 64, 128, 256, 512, 16]
    ins, outs, specs = make_batch(zh, torch, sides, seed=8, col_scale=8)
    saved = {s: zh.get_variant(s) for s in set(sides)}
    tuned = zh.autotune_batch(specs, iters=2)
    print("[autotune_batch]", {s: (e["choice"], e["times_us"]) for s, e in tuned.items()}, flush=True)
    for s, e in tuned.items():
        best = min(range(zh.TUNE_CANDIDATES), key=lambda v: e["times_us"][v])
        assert e["times_us"][e["choice"]] == e["times_us"][best], (s, e)
    # the tuner's own launches wrote the squares, and the tuned table encodes them the same
    for t_in, t_out, s in zip(ins, outs, sides):
        assert torch.equal(t_out, plain(zh, torch, t_in, s)), s
    for o in outs:
        o.zero_()
    zh.encode_batch(specs)
    torch.cuda.synchronize()
    for t_in, t_out, s in zip(ins, outs, sides):
        assert torch.equal(t_out, plain(zh, torch, t_in, s)), s
    for s, v in saved.items():
        zh.set_variant(s, v)


def test_batch_argument_errors(zh, torch):
    # every call here is rejected on the host before any launch; the buffers
    # are still sized for the shapes named, so nothing could write past them
    t_in = torch.empty((8192, 8), dtype=torch.int32, device="cuda")
    t_out = torch.empty((16384, 8), dtype=torch.int32, device="cuda")
    with pytest.raises(zh.ZodaError):
        zh.encode_batch([(t_in.data_ptr(), t_out.data_ptr(), 8192, 8, 8192)])  # side above 4096
    with pytest.raises(zh.ZodaError):
        zh.encode_batch([(t_in.data_ptr(), t_out.data_ptr(), 8, 8, 0)])        # n_kn = n_k
    with pytest.raises(zh.ZodaError):
        zh.encode_batch([(t_in.data_ptr(), t_out.data_ptr(), 8, 8, 9)])        # n_kn = 4 n_k
    with pytest.raises(zh.ZodaError):
        zh.encode_batch([(t_in.data_ptr(), t_out.data_ptr(), 0, 8, 8)])        # k = 0
    with pytest.raises(zh.ZodaError):
        zh.encode_batch([])
