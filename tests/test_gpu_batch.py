"""The batched advance (one launch per pass and kernel variant over many
instances, optionally replayed from a hipGraph) reproduces the
single-instance advance of every member bit for bit."""

import numpy as np
import pytest
import torch

from oracle import model as M

pytestmark = pytest.mark.gpu

# instance set: several shapes selecting different kernel variants (register
# and LDS z plans, register and generic x/y FFTs), alphas and noise scales
SHAPES = [((32, 32, 24), 1.8, 0.3), ((64, 48, 40), 1.6, 0.25),
          ((48, 40, 30), 1.7, 0.2), ((64, 64, 32), 1.9, 0.35),
          ((32, 32, 24), 1.5, 0.15)]


def _instances(seed=0, defer=False):
    from steam_amd import simulate as S
    from steam_amd.batch import Instance
    kern = S._turbulon_envelope(1, .5, .5, .5, as_tensor=True)
    rng = np.random.default_rng(seed)
    out = []
    for i, (shape, alpha, c) in enumerate(SHAPES):
        flux = torch.from_numpy(np.abs(1 + 0.2 * rng.standard_normal(shape))
                                .astype(np.float32)).cuda()
        key = M.class_key(np.random.SeedSequence(100 + i))
        region = S.NoiseRegion(tuple(np.uint64(k) for k in key), (0, 0, 0), shape)
        out.append(Instance(flux, kern, region, c, alpha=alpha,
                            defer_rescale=defer))
    return out


def _single(inst, op, spec, region=None):
    from steam_amd import simulate as S
    f = inst.flux.clone()
    S.advance_flux_device(f, inst.region if region is None else region, inst.kernel,
                          inst.flux_noise_scale, 1, (1, 1, 1), alpha=inst.alpha,
                          op=op, spec=spec)
    return f


def test_batched_advance_equals_single_instance_advance():
    from steam_amd.batch import BatchedAdvance
    insts = _instances()
    batch = BatchedAdvance(insts, autotune=False)
    refs = [_single(inst, op, sp) for inst, op, sp in zip(insts, batch.ops, batch.specs)]
    batch.advance()
    torch.cuda.synchronize()
    for inst, ref in zip(insts, refs):
        np.testing.assert_array_equal(inst.flux.cpu().numpy(), ref.cpu().numpy())
    assert batch.launches_per_advance < 8 * len(insts)


def test_batched_ledgers_match_oracle_sums():
    from steam_amd.batch import BatchedAdvance
    insts = _instances(seed=1)
    start = [inst.flux.cpu().numpy() for inst in insts]
    batch = BatchedAdvance(insts, autotune=False)
    batch.advance()
    for inst, f0, led in zip(insts, start, batch.ledgers()):
        kern = inst.kernel.cpu().numpy()
        fo, so, lo, do = M.advance_flux(f0, inst.region.key, (0, 0, 0), f0.shape, kern,
                                        inst.flux_noise_scale, 1, (1, 1, 1),
                                        alpha=inst.alpha)
        assert abs(led.entering.total / lo['entering'][0] - 1) < 1e-12
        assert abs(led.realized.total / lo['realized'][0] - 1) < 1e-5
        f = inst.flux.cpu().numpy().astype(np.float64)
        assert np.linalg.norm(f - fo) / np.linalg.norm(fo) < 1e-5


def test_graph_replays_match_eager_and_follow_new_keys():
    """Replays with eager work and synchronisations between them; every
    replay is checked against the eager batched result."""
    from steam_amd import simulate as S
    from steam_amd.batch import BatchedAdvance
    insts = _instances(seed=2)
    twins = _instances(seed=2)
    g_batch = BatchedAdvance(insts, autotune=False)
    e_batch = BatchedAdvance(twins, autotune=False)
    graph = g_batch.capture()  # capture() itself runs one warm advance
    e_batch.advance()
    torch.cuda.synchronize()
    for rep in range(3):
        regions = [S.NoiseRegion(tuple(np.uint64(k) for k in
                                       M.class_key(np.random.SeedSequence(7 * rep + i))),
                                 (0, 0, 0), tuple(inst.flux.shape))
                   for i, inst in enumerate(insts)]
        torch.cuda.synchronize()
        g_batch.set_regions(regions)
        e_batch.set_regions(regions)
        graph.replay()
        e_batch.advance()
        torch.cuda.synchronize()
        junk = torch.randn(1 << 20, device='cuda').sum().item()  # eager work between replays
        assert np.isfinite(junk)
        for a, b in zip(insts, twins):
            np.testing.assert_array_equal(a.flux.cpu().numpy(), b.flux.cpu().numpy())


def test_replay_then_new_keys_without_synchronising():
    """set_regions right after a replay, with no synchronisation in between:
    each replay must still run with the keys set before it (the staging
    copy is not rewritten until the replay's upload has read it)."""
    from steam_amd import simulate as S
    from steam_amd.batch import BatchedAdvance
    insts = _instances(seed=4)
    twins = _instances(seed=4)
    g_batch = BatchedAdvance(insts, autotune=False)
    e_batch = BatchedAdvance(twins, autotune=False)
    graph = g_batch.capture()
    e_batch.advance()
    torch.cuda.synchronize()

    def regions(rep):
        return [S.NoiseRegion(tuple(np.uint64(k) for k in
                                    M.class_key(np.random.SeedSequence(11 * rep + i))),
                              (0, 0, 0), tuple(inst.flux.shape))
                for i, inst in enumerate(insts)]
    for rep in range(4):
        g_batch.set_regions(regions(rep))
        graph.replay()
    for rep in range(4):
        e_batch.set_regions(regions(rep))
        e_batch.advance()
        torch.cuda.synchronize()
    torch.cuda.synchronize()
    for a, b in zip(insts, twins):
        np.testing.assert_array_equal(a.flux.cpu().numpy(), b.flux.cpu().numpy())


def test_deferred_batch_leaves_rescale_pending():
    from steam_amd.batch import BatchedAdvance
    insts = _instances(seed=3, defer=True)
    twins = _instances(seed=3, defer=False)
    b1 = BatchedAdvance(insts, autotune=False)
    b2 = BatchedAdvance(twins, autotune=False)
    b1.advance()
    b2.advance()
    for a, b, led in zip(insts, twins, b1.ledgers()):
        scaled = (a.flux * np.float32(led.rescale)).cpu().numpy()
        np.testing.assert_array_equal(scaled, b.flux.cpu().numpy())


def test_autotune_picks_a_timed_candidate():
    from steam_amd.batch import autotune_nfft, nfft_candidates
    cands = nfft_candidates(40, 13, max_out=3)
    assert len(cands) == 3 and all(n % 2 == 0 and n >= 13 for n in cands)
    best, times = autotune_nfft((32, 32, 40), (13, 13, 13), top=3, reps=2, table={})
    assert best in cands and set(times) == set(cands)
    assert times[best] == min(times.values())


def test_batch_rejects_bad_items():
    from steam_amd.batch import BatchedAdvance, Instance
    insts = _instances()
    with pytest.raises(ValueError):
        BatchedAdvance([])
    with pytest.raises(ValueError):
        Instance(insts[0].flux.cpu(), insts[0].kernel, insts[0].region, 0.2)
    bad = _instances()
    bad[0].window = (0, 999, 0, 4)
    with pytest.raises(ValueError):
        BatchedAdvance(bad, autotune=False)


def test_pipelined_batch_equals_the_batched_advance():
    """Noise generated a step ahead on a side stream into two buffer sets,
    each step's rescale left pending and applied on the next step's read
    (eager steps, then graph replays with new keys each step and no
    synchronisation between them): the batched advance's result bit for
    bit."""
    from steam_amd import simulate as S
    from steam_amd.batch import BatchedAdvance, PipelinedBatch
    insts = _instances(seed=5)
    twins = _instances(seed=5)
    pipe = PipelinedBatch(insts, autotune=False)
    plain = BatchedAdvance(twins, autotune=False)

    def regions(step):
        # step 0: the instances' own regions
        if step == 0:
            return [inst.region for inst in twins]
        return [S.NoiseRegion(tuple(np.uint64(k) for k in
                                    M.class_key(np.random.SeedSequence(1000 * step + i))),
                              (0, 0, 0), tuple(inst.flux.shape))
                for i, inst in enumerate(twins)]
    steps = 7
    for step in range(steps):
        pipe.set_regions(regions(step + 1))
        if step < 2:
            pipe.advance()
        else:
            if pipe.graphs is None:
                pipe.capture()
            pipe.replay()
    pipe.finish()           # the last step's rescale, left pending
    torch.cuda.synchronize()
    for step in range(steps):
        plain.set_regions(regions(step))
        plain.advance()
    torch.cuda.synchronize()
    for a, b in zip(insts, twins):
        np.testing.assert_array_equal(a.flux.cpu().numpy(), b.flux.cpu().numpy())
    for la, lb in zip(pipe.ledgers(), plain.ledgers()):
        assert la.realized == lb.realized and la.n_clipped == lb.n_clipped


def test_crops_store_a_box_of_the_whole_grid_bit_for_bit():
    """An apply-mode instance with a unit entering flux and a register z plan
    (a served tile) stores only a box of its advanced flux, dense in a
    caller tensor: the box equals the same box of its whole-grid advance bit
    for bit, odd z corners and odd level counts included.  A measure-mode
    instance and a box outside the grid are refused."""
    from steam_amd import simulate as S
    from steam_amd.batch import BatchedAdvance, Instance
    kern = S._turbulon_envelope(1, .5, .5, .5, as_tensor=True)
    shapes = [(64, 48, 40), (48, 40, 31), (64, 64, 32)]
    boxes = [((5, 7, 3), (40, 30, 33)), ((0, 0, 1), (48, 40, 30)), ((10, 20, 0), (8, 8, 32))]
    regions = []
    for i, shape in enumerate(shapes):
        key = M.class_key(np.random.SeedSequence(300 + i))
        regions.append(S.NoiseRegion(tuple(np.uint64(k) for k in key), (0, 0, 0), shape))
    # the ledgers: a measure-mode advance of each grid from a unit flux
    measure = [Instance(torch.ones(sh, device='cuda'), kern, r, 0.3, alpha=1.8)
               for sh, r in zip(shapes, regions)]
    mb = BatchedAdvance(measure, autotune=False)
    mb.advance()
    ledgers = mb.ledgers()
    with pytest.raises(ValueError):
        mb.set_crops([(torch.empty(shapes[0], device='cuda'), (0, 0, 0)), None, None])

    def apply_batch():
        insts = [Instance(torch.empty(sh, device='cuda'), kern, r, 0.3, alpha=1.8,
                          ledger=led, unit_flux=True)
                 for sh, r, led in zip(shapes, regions, ledgers)]
        return insts, BatchedAdvance(insts, autotune=False)

    insts, b = apply_batch()
    b.advance()
    torch.cuda.synchronize()
    whole = [inst.flux.cpu().numpy() for inst in insts]
    outs = [torch.full(n, -1.0, device='cuda') for _, n in boxes]
    b.set_crops([(o, org) for o, (org, _) in zip(outs, boxes)])
    for inst in insts:
        inst.flux.fill_(-2.0)
    b.advance()
    torch.cuda.synchronize()
    for i, ((o, n), out) in enumerate(zip(boxes, outs)):
        ref = whole[i][o[0]:o[0] + n[0], o[1]:o[1] + n[1], o[2]:o[2] + n[2]]
        print(f"instance {i}: box {o} + {n}", flush=True)
        np.testing.assert_array_equal(out.cpu().numpy(), ref)
        # the grid itself is not written while a crop is set
        assert np.all(insts[i].flux.cpu().numpy() == -2.0)
    # back to whole grids
    b.set_crops([None] * len(insts))
    b.advance()
    torch.cuda.synchronize()
    for inst, ref in zip(insts, whole):
        np.testing.assert_array_equal(inst.flux.cpu().numpy(), ref)
    with pytest.raises(ValueError):
        b.set_crops([(torch.empty((8, 8, 8), device='cuda'), (60, 0, 0)), None, None])


def test_members_on_bluestein_axes_equal_their_single_advance():
    """Members whose x or y length is prime (the Bluestein form of the x/y
    passes, 512-thread workgroups) beside an ordinary member: the batched
    advance equals each member's single advance bit for bit."""
    from steam_amd import simulate as S
    from steam_amd.batch import BatchedAdvance, Instance
    kern = S._turbulon_envelope(1, .5, .5, .5, as_tensor=True)
    rng = np.random.default_rng(5)
    insts = []
    for i, shape in enumerate([(40, 131, 12), (137, 24, 10), (32, 32, 24)]):
        flux = torch.from_numpy(np.abs(1 + 0.2 * rng.standard_normal(shape))
                                .astype(np.float32)).cuda()
        key = M.class_key(np.random.SeedSequence(500 + i))
        region = S.NoiseRegion(tuple(np.uint64(k) for k in key), (0, 0, 0), shape)
        insts.append(Instance(flux, kern, region, 0.3, alpha=1.8))
    batch = BatchedAdvance(insts, autotune=False)
    refs = [_single(inst, op, sp) for inst, op, sp in zip(insts, batch.ops, batch.specs)]
    batch.advance()
    torch.cuda.synchronize()
    for inst, ref in zip(insts, refs):
        np.testing.assert_array_equal(inst.flux.cpu().numpy(), ref.cpu().numpy())
