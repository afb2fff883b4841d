"""World-keyed tile serving with continuous batching.

One class advance of a world field (e.g. 2048 x 2048 x 256), served as tiles.
The class's global means (entering flux, |noise|, realized flux) are
measured once per world by a tiled measure pass (the streamed cascade's
measure/reduce passes, steam/streaming.py:850-1447, restricted to the flux).
Every tile request then runs the apply-only advance
(steam/simulate.py:2066-2228 with ``ledger=``) on its tile plus a halo of
the kernel half-width: the keyed noise is a function of world coordinates
(steam/streaming.py:822-847), so the tile's interior equals the monolithic
solve restricted to it, seams included.

The server admits requests as they arrive, against a device-memory budget
and a pool of streams.  Waiting requests of the same padded tile shape are
merged into one batched advance (one launch per pass for the group,
steam_batch_*), and groups run concurrently on their own streams (their
kernels overlap on the device); new arrivals are admitted as soon as a
stream and memory free up.  Kernel spectra are cached per padded tile shape,
and each (shape, stream) keeps its slot buffers and batch descriptors, so
nothing is allocated per request; finished tiles are kept in a
device-resident LRU cache so a request for a recently served tile (a
neighbour walk coming back) is answered without recomputation.  Latency is
arrival to completion of the tile on the device.
"""

import time
from collections import OrderedDict, deque

import numpy as np
import torch

from . import _lib
from .ledger import FluxAdvanceLedger, MeanReduction
from .simulate import NoiseRegion, _turbulon_envelope, advance_flux_device, class_key
from .utils import conv_operator, kernel_key


def fft_length(m):
    """Transform length of a padded tile axis of at least m cells: the
    smallest length with a register x/y FFT path (a power of two 64..2048 or
    a compiled mixed-radix length, e.g. 80, 144, 288, 400) when it is within
    30% of m, else the smallest 2^a 3^b 5^c (the LDS-staged passes)."""
    from .utils import register_fft_lengths
    for n in register_fft_lengths():
        if n >= m:
            if n <= 1.3 * m:
                return n
            break
    n = m
    while True:
        r = n
        for f in (2, 3, 5):
            while r % f == 0:
                r //= f
        if r == 1:
            return n
        n += 1


class WorldClass:
    """One flux class of a world: keyed noise, parameters, kernel, halo."""

    def __init__(self, world_shape, seed=0, flux_noise_scale=0.1 ** (1 / 1.8),
                 alpha=1.8, k=1.0, spacing=(0.5, 0.5, 0.5), support_factor=3):
        self.world_shape = tuple(int(n) for n in world_shape)
        self.key = class_key(np.random.SeedSequence(seed).spawn(1)[0])
        self.flux_noise_scale = float(flux_noise_scale)
        self.alpha = float(alpha)
        self.kernel = _turbulon_envelope(k, *spacing, support_factor=support_factor,
                                         as_tensor=True)
        # halo per axis = kernel half-width (+1 as the streamed driver keeps,
        # steam/streaming.py:672-796)
        self.halo = tuple(int(n) // 2 + 1 for n in self.kernel.shape)
        self.ledger = None

    def padded_box(self, origin, shape):
        """(padded origin, padded shape, interior offset) of a tile: x/y
        grow by at least the halo on both sides (periodic world) up to an
        FFT-friendly length, z by the halo inside the world only (z is not
        periodic), kept even.  A wider margin than the halo leaves the
        interior exact."""
        (x0, y0, z0), (tx, ty, tz) = origin, shape
        hx, hy, hz = self.halo
        nz = self.world_shape[2]
        px, py = fft_length(tx + 2 * hx), fft_length(ty + 2 * hy)
        zlo = max(0, z0 - hz)
        zhi = min(nz, z0 + tz + hz)
        if (zhi - zlo) % 2:
            if zhi < nz:
                zhi += 1
            elif zlo > 0:
                zlo -= 1
        return ((x0 - hx, y0 - hy, zlo), (px, py, zhi - zlo), (hx, hy, z0 - zlo))

    def region(self, padded_origin):
        return NoiseRegion(self.key, tuple(int(v) for v in padded_origin),
                           self.world_shape)

    def advance_tile(self, origin, shape, ledger=None, stream=None):
        """Flux after the class advance on one tile (device tensor).  With a
        ledger: apply-only (the world's means); without: measure mode over
        the tile interior, returning the tile's ledger too."""
        po, ps, (ix, iy, iz) = self.padded_box(origin, shape)
        flux = torch.empty(ps, dtype=torch.float32, device='cuda')
        op = conv_operator(ps, tuple(self.kernel.shape))
        spec = op.spectrum(self.kernel, key=kernel_key(self.kernel))
        window = None if ledger is not None else (ix, ix + shape[0], iy, iy + shape[1])
        if ledger is None and (iz != 0 or ps[2] != shape[2]):
            raise ValueError("measure tiles must span the whole z column")
        led, _ = advance_flux_device(flux, self.region(po), self.kernel,
                                     self.flux_noise_scale, 1, (1, 1, 1),
                                     alpha=self.alpha, ledger=ledger, op=op,
                                     spec=spec, window=window, unit_flux=True)
        out = flux[ix:ix + shape[0], iy:iy + shape[1], iz:iz + shape[2]]
        return out, led

    def measure(self, tile=(512, 512)):
        """World ledger from a tiled measure pass (x/y tiles, whole z)."""
        nx, ny, nz = self.world_shape
        tx, ty = min(tile[0], nx), min(tile[1], ny)
        ent = nab = rea = None
        clipped = 0
        for x0 in range(0, nx, tx):
            for y0 in range(0, ny, ty):
                shape = (min(tx, nx - x0), min(ty, ny - y0), nz)
                _, led = self.advance_tile((x0, y0, 0), shape)
                ent = led.entering if ent is None else ent + led.entering
                nab = led.noise_abs if nab is None else nab + led.noise_abs
                rea = led.realized if rea is None else rea + led.realized
                clipped += led.n_clipped  # per call, halo cells included
        self.ledger = FluxAdvanceLedger(ent, nab, rea, clipped)
        return self.ledger

    def monolithic(self):
        """The whole world advanced in one piece (measure mode)."""
        flux = torch.ones(self.world_shape, dtype=torch.float32, device='cuda')
        led, _ = advance_flux_device(flux, NoiseRegion(self.key, (0, 0, 0),
                                                       self.world_shape),
                                     self.kernel, self.flux_noise_scale, 1, (1, 1, 1),
                                     alpha=self.alpha)
        return flux, led


def make_trace(n, rate, world_shape, sizes=((64, 64, 64), (128, 128, 128),
                                            (256, 256, 256), (384, 384, 192)),
               p_neighbor=0.7, seed=0):
    """Poisson arrivals at ``rate`` requests/s; tile size uniform over
    ``sizes``; with probability ``p_neighbor`` the origin is a periodic
    neighbour (one tile extent in x or y) of the previous request's."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    prev = None
    nx, ny, nz = world_shape
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        size = tuple(sizes[int(rng.integers(len(sizes)))])
        size = (min(size[0], nx), min(size[1], ny), min(size[2], nz))
        if prev is not None and rng.random() < p_neighbor:
            (px, py, pz), psz = prev
            axis = int(rng.integers(2))
            step = int(rng.choice((-1, 1)))
            x = (px + (step * psz[0] if axis == 0 else 0)) % nx
            y = (py + (step * psz[1] if axis == 1 else 0)) % ny
            z = min(pz, nz - size[2])
        else:
            x = int(rng.integers(nx))
            y = int(rng.integers(ny))
            z = int(rng.integers(nz - size[2] + 1))
        origin = (int(x), int(y), int(z))
        out.append({"id": i, "arrival": t, "origin": origin, "shape": size})
        prev = (origin, size)
    return out


class _InFlight:
    __slots__ = ("reqs", "stream", "event", "outs", "pool", "bytes")


class _SlotPool:
    """The device buffers of one padded tile shape on one stream: ``slots``
    flux buffers (views of one tensor; each advance starts from a unit flux
    it does not read, so they are never filled) and workspaces, and a
    batched advance over the first c slots for each group
    size c, created on first use and kept (their descriptors are refreshed
    per group from pinned memory, nothing is allocated per request)."""

    def __init__(self, world, padded_shape, slots):
        self.world = world
        self.shape = tuple(padded_shape)
        self.flux = torch.ones((slots,) + self.shape, dtype=torch.float32, device='cuda')
        op = conv_operator(self.shape, tuple(world.kernel.shape))
        self.workspaces = [torch.empty(op.workspace_bytes, dtype=torch.uint8, device='cuda')
                           for _ in range(slots)]
        self.batches = {}
        # whether the batches store each tile's interior straight into its
        # result (a register z plan), else it is copied out of the slot
        self.crops = True

    @property
    def nbytes(self):
        return self.flux.numel() * 4 + sum(w.numel() for w in self.workspaces)

    def batch(self, count):
        from .batch import BatchedAdvance, Instance
        b = self.batches.get(count)
        if b is None:
            w = self.world
            insts = [Instance(self.flux[i], w.kernel, w.region((0, 0, 0)), w.flux_noise_scale,
                              alpha=w.alpha, ledger=w.ledger, unit_flux=True)
                     for i in range(count)]
            b = BatchedAdvance(insts, autotune=False, workspaces=self.workspaces[:count])
            self.batches[count] = b
        return b


class TileServer:
    """Continuous-batching tile server over one WorldClass.

    Requests wait in arrival order.  Whenever a stream is free, the oldest
    waiting request is admitted together with up to ``max_batch - 1`` later
    waiting requests of the same padded tile shape (within the device-memory
    budget): their apply-only advances run as one batched advance (one
    launch per pass for the whole group, steam_batch_*) on that stream,
    while other groups run on the other streams.  A tile's result does not
    depend on its group: the batched advance of an instance equals its
    single advance bit for bit."""

    def __init__(self, world, budget_bytes=8 << 30, streams=4, cache_bytes=2 << 30,
                 max_batch=8, pool_bytes=2 << 30):
        if world.ledger is None:
            raise ValueError("measure the world ledger first (WorldClass.measure)")
        _lib.require_gpu('cuda')
        self.world = world
        self.budget = int(budget_bytes)
        self.streams = [torch.cuda.Stream() for _ in range(max(1, streams))]
        self.max_batch = max(1, int(max_batch))
        self.pool_bytes = int(pool_bytes)
        self._slot_counts = {}
        self.cache_bytes = int(cache_bytes)
        self.cache = OrderedDict()
        self.cache_used = 0
        self.pools = {}
        self.stats = {"cache_hits": 0, "spectra": 0, "max_inflight": 0,
                      "max_bytes": 0, "groups": 0, "batched_requests": 0, "max_group": 0}

    @staticmethod
    def request_bytes(world, shape):
        _, ps, _ = world.padded_box((0, 0, 0), shape)
        return 4 * int(np.prod(ps)) + 4 * int(np.prod(shape))

    def slots(self, ps):
        """Group size cap of a padded shape: max_batch, or fewer when the
        slot buffers of one stream would pass pool_bytes."""
        ps = tuple(ps)
        n = self._slot_counts.get(ps)
        if n is None:
            op = conv_operator(ps, tuple(self.world.kernel.shape))
            per_slot = 4 * int(np.prod(ps)) + int(op.workspace_bytes)
            n = max(1, min(self.max_batch, self.pool_bytes // per_slot))
            self._slot_counts[ps] = n
        return n

    def _pool(self, ps, stream):
        key = (tuple(ps), stream.cuda_stream)
        pool = self.pools.get(key)
        if pool is None:
            w = self.world
            op = conv_operator(ps, tuple(w.kernel.shape))
            kkey = kernel_key(w.kernel)
            if kkey not in op._spectra:
                op.spectrum(w.kernel, key=kkey)
                torch.cuda.current_stream().synchronize()
                self.stats["spectra"] += 1
            pool = _SlotPool(w, ps, self.slots(ps))
            self.pools[key] = pool
        return pool

    def _launch(self, reqs, stream):
        w = self.world
        boxes = [w.padded_box(r["origin"], r["shape"]) for r in reqs]
        pool = self._pool(boxes[0][1], stream)
        b = pool.batch(len(reqs))
        with torch.cuda.stream(stream):
            # the advance starts from a unit flux it does not read (no fill);
            # the stream is free: the batch's previous upload has been read
            b.set_regions([w.region(po) for po, _, _ in boxes])
            outs = None
            if pool.crops:
                # the inverse z pass stores each tile's interior in its result
                outs = [torch.empty(tuple(r["shape"]), dtype=torch.float32, device='cuda')
                        for r in reqs]
                try:
                    b.set_crops([(o, inner) for o, (_, _, inner) in zip(outs, boxes)])
                except ValueError:
                    pool.crops = False
                    b.set_crops([None] * len(reqs))
                    outs = None
            b.advance(stream)
            if outs is None:
                outs = []
                for i, (r, (_, _, (ix, iy, iz))) in enumerate(zip(reqs, boxes)):
                    tx, ty, tz = r["shape"]
                    outs.append(pool.flux[i, ix:ix + tx, iy:iy + ty, iz:iz + tz].contiguous())
            ev = torch.cuda.Event()
            ev.record(stream)
        return pool, outs, ev

    def _cache_put(self, key, t):
        nb = t.numel() * 4
        if nb > self.cache_bytes:
            return
        while self.cache_used + nb > self.cache_bytes and self.cache:
            _, old = self.cache.popitem(last=False)
            self.cache_used -= old.numel() * 4
        self.cache[key] = t
        self.cache_used += nb

    def serve(self, trace, keep_results=False):
        """Run a request trace in real time.  Returns (latencies_s,
        wall_s, voxels, results) with results[id] = device tile when
        ``keep_results`` (True: every request; a set: those ids)."""
        wanted = keep_results if isinstance(keep_results, (set, frozenset)) else None
        keep_results = bool(keep_results)

        def keep(rid):
            return keep_results and (wanted is None or rid in wanted)

        w = self.world
        pending = deque(sorted(trace, key=lambda r: r["arrival"]))
        queue = []
        inflight = []
        free_streams = list(self.streams)
        used = 0
        lat = {}
        results = {}
        voxels = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        while pending or queue or inflight:
            now = time.perf_counter() - t0
            while pending and pending[0]["arrival"] <= now:
                queue.append(pending.popleft())
            # completions first: they free streams and memory
            still = []
            for f in inflight:
                if f.event.query():
                    done = time.perf_counter() - t0
                    for r, out in zip(f.reqs, f.outs):
                        lat[r["id"]] = done - r["arrival"]
                        voxels += out.numel()
                        self._cache_put((r["origin"], r["shape"]), out)
                        if keep(r["id"]):
                            results[r["id"]] = out
                    used -= f.bytes
                    free_streams.append(f.stream)
                else:
                    still.append(f)
            inflight = still
            # exact repeats of a finished tile are answered from the cache
            rest = []
            for req in queue:
                key = (req["origin"], req["shape"])
                hit = self.cache.get(key)
                if hit is None:
                    rest.append(req)
                    continue
                self.cache.move_to_end(key)
                self.stats["cache_hits"] += 1
                lat[req["id"]] = (time.perf_counter() - t0) - req["arrival"]
                voxels += hit.numel()
                if keep(req["id"]):
                    results[req["id"]] = hit
            queue = rest
            # admission: the oldest waiting request and the later waiting
            # requests of its padded shape, while a stream and memory are free
            while queue and free_streams:
                head = queue[0]
                ps = w.padded_box(head["origin"], head["shape"])[1]
                cap = self.slots(ps)
                group, nb = [], 0
                for req in queue:
                    if len(group) == cap:
                        break
                    if w.padded_box(req["origin"], req["shape"])[1] != ps:
                        continue
                    rb = self.request_bytes(w, req["shape"])
                    if group and used + nb + rb > self.budget:
                        break
                    group.append(req)
                    nb += rb
                if inflight and used + nb > self.budget:
                    break
                ids = {r["id"] for r in group}
                queue = [r for r in queue if r["id"] not in ids]
                s = free_streams.pop()
                f = _InFlight()
                f.reqs, f.stream, f.bytes = group, s, nb
                f.pool, f.outs, f.event = self._launch(group, s)
                used += nb
                inflight.append(f)
                self.stats["groups"] += 1
                self.stats["batched_requests"] += len(group)
                self.stats["max_group"] = max(self.stats["max_group"], len(group))
                self.stats["max_inflight"] = max(self.stats["max_inflight"],
                                                 sum(len(x.reqs) for x in inflight))
                self.stats["max_bytes"] = max(self.stats["max_bytes"], used)
            if not queue and not inflight and pending:
                wait = pending[0]["arrival"] - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 1e-3))
        wall = time.perf_counter() - t0
        lats = np.array([lat[r["id"]] for r in trace])
        return lats, wall, voxels, results
