"""Batched flux advance and the FFT schedule autotuner.

The reference advances one realisation or tile per call
(steam/simulate.py:2066-2228).  ``BatchedAdvance`` advances many independent
instances of different size, stability index and noise scale together: one
launch per pass and compiled kernel variant over all instances
(``steam_batch_advance``), optionally captured once into a hipGraph and
replayed.  Each instance's result is the one the single-instance advance
(``simulate.advance_flux_device``) gives with the same plan.

``autotune_nfft`` times the planner's leading z transform lengths for a
(field, kernel) shape pair and keeps the fastest; the choice fixes the radix
schedule of the z passes (register plan or LDS stages) and the number of
spectral planes of the x/y passes.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .ledger import FluxAdvanceLedger
from .simulate import class_scale_shift, noise_params
from .utils import conv_operator, kernel_key

_TUNED = {}


def nfft_candidates(nz, kz, max_out=8):
    """Usable z transform lengths for (nz, kz), best model cost first."""
    lib = _lib.load()
    out = (ctypes.c_int64 * max_out)()
    n = ctypes.c_int32()
    _lib.check(lib.steam_conv_nfft_candidates(int(nz), int(kz), out, max_out,
                                              ctypes.byref(n)),
               "steam_conv_nfft_candidates")
    return [int(out[i]) for i in range(n.value)]


def autotune_nfft(field_shape, kernel_shape, top=3, reps=10, trials=3, table=None):
    """Fastest z transform length among the ``top`` model candidates, timed
    with HIP events on the convolution operator (cached per shape pair): the
    best of ``trials`` runs of ``reps`` transforms each, the candidates taken
    in turn within a trial, so a transient stall does not decide the choice
    for a small field.

    Returns (n_fft, {n_fft: ms}).
    """
    table = _TUNED if table is None else table
    key = (tuple(int(n) for n in field_shape), tuple(int(n) for n in kernel_shape))
    if key in table:
        return table[key]
    _lib.require_gpu('cuda')
    cands = nfft_candidates(field_shape[2], kernel_shape[2], max_out=top)
    gen = torch.Generator(device='cuda')
    gen.manual_seed(0)
    field = torch.randn(key[0], device='cuda', generator=gen)
    kern = torch.randn(key[1], device='cuda', generator=gen)
    out = torch.empty_like(field)
    ops = {}
    for n in cands:
        op = conv_operator(key[0], key[1], n_fft=n)
        spec = op.spectrum(kern)
        op.execute(field, out, spec)
        ops[n] = (op, spec)
    times = {}
    for _ in range(trials):
        for n, (op, spec) in ops.items():
            start = torch.cuda.Event(enable_timing=True)
            stop = torch.cuda.Event(enable_timing=True)
            start.record()
            for _ in range(reps):
                op.execute(field, out, spec)
            stop.record()
            stop.synchronize()
            t = start.elapsed_time(stop) / reps
            times[n] = min(times.get(n, t), t)
    best = min(times, key=times.get)
    table[key] = (best, times)
    return best, times


class Instance:
    """One member of a batch: its flux state (device tensor, advanced in
    place), kernel (device tensor), noise region and class parameters."""

    def __init__(self, flux, kernel, region, flux_noise_scale,
                 n_scale_classes_per_dyad=1, sparsity_factors=(1, 1, 1),
                 alpha=None, n_zero=0, zero_bottom=False, zero_top=False,
                 window=None, defer_rescale=False, chain_pending=False, ledger=None,
                 unit_flux=False):
        if not (isinstance(flux, torch.Tensor) and flux.is_cuda
                and flux.dtype == torch.float32 and flux.is_contiguous()):
            raise ValueError("batched instances need contiguous float32 device "
                             "tensors for the flux")
        self.flux = flux
        self.kernel = kernel.contiguous()
        self.region = region
        self.flux_noise_scale = flux_noise_scale
        self.n_scale_classes_per_dyad = n_scale_classes_per_dyad
        self.sparsity_factors = tuple(sparsity_factors)
        self.alpha = alpha
        self.n_zero = n_zero
        self.zero_bottom = zero_bottom
        self.zero_top = zero_top
        self.window = window
        self.defer_rescale = defer_rescale
        self.chain_pending = chain_pending
        # a recorded FluxAdvanceLedger: the instance advances apply-only
        self.ledger = ledger
        # the entering flux is 1 everywhere; the flux tensor is only written
        self.unit_flux = unit_flux

    def native_noise(self, region=None):
        a, pcs, shift = class_scale_shift(self.flux_noise_scale,
                                          self.n_scale_classes_per_dyad, self.alpha)
        return noise_params(self.region if region is None else region,
                            self.sparsity_factors, a, pcs, shift, self.n_zero,
                            self.zero_bottom, self.zero_top)


class BatchedAdvance:
    """Measure-mode class advance of many instances per call.

    ``autotune`` chooses each distinct shape pair's z transform length by
    timing (``autotune_nfft``); otherwise the planner's model choice is used.
    Every instance owns its workspace, so its ledger can be read back after
    any advance (``ledgers``).
    """

    def __init__(self, instances, autotune=True, noise_buffers=None, workspaces=None):
        lib = _lib.load()
        self.instances = list(instances)
        if not self.instances:
            raise ValueError("empty batch")
        if noise_buffers is not None:
            noise_buffers = list(noise_buffers)
            if len(noise_buffers) != len(self.instances) or any(
                    tuple(b.shape) != tuple(inst.flux.shape) or b.dtype != torch.float32
                    or not b.is_cuda or not b.is_contiguous()
                    for b, inst in zip(noise_buffers, self.instances)):
                raise ValueError("one contiguous float32 device noise buffer of the "
                                 "instance's shape per instance expected")
        self.noise_buffers = noise_buffers
        self.ops, self.specs, self.workspaces = [], [], []
        self._ledgers = []
        self.n_fft = []
        items = (_lib.BatchItem * len(self.instances))()
        for i, inst in enumerate(self.instances):
            shape, kshape = tuple(inst.flux.shape), tuple(inst.kernel.shape)
            n_fft = autotune_nfft(shape, kshape)[0] if autotune else 0
            op = conv_operator(shape, kshape, n_fft=n_fft)
            spec = op.spectrum(inst.kernel, key=kernel_key(inst.kernel))
            if workspaces is not None:
                ws = workspaces[i]
                if ws.numel() < op.workspace_bytes:
                    raise ValueError("a shared workspace is smaller than the plan needs")
            else:
                ws = torch.empty(op.workspace_bytes, dtype=torch.uint8, device='cuda')
            self.ops.append(op)
            self.specs.append(spec)
            self.workspaces.append(ws)
            self.n_fft.append(op.n_fft)
            it = items[i]
            it.plan = op.plan
            it.spectrum_dev = spec.data_ptr()
            it.workspace_dev = ws.data_ptr()
            it.flux_dev = inst.flux.data_ptr()
            it.noise_dev = None if noise_buffers is None else noise_buffers[i].data_ptr()
            it.sk_dev = None
            it.noise = inst.native_noise()
            if inst.window is not None:
                for a in range(4):
                    it.window[a] = int(inst.window[a])
            it.sigma_in = 1.0
            it.flags = ((_lib.DEFER_RESCALE if inst.defer_rescale else 0)
                        | (_lib.CHAIN_PENDING if inst.chain_pending else 0)
                        | (_lib.UNIT_FLUX if getattr(inst, "unit_flux", False) else 0))
            if inst.ledger is not None:
                native = inst.ledger.to_native()
                self._ledgers.append(native)
                it.ledger = ctypes.pointer(native)
        self._items = items
        handle = ctypes.c_void_p()
        _lib.check(lib.steam_batch_create(items, len(self.instances),
                                          ctypes.byref(handle)),
                   "steam_batch_create")
        self.handle = handle
        launches = ctypes.c_int32()
        desc = ctypes.c_int64()
        _lib.check(lib.steam_batch_info(handle, ctypes.byref(launches),
                                        ctypes.byref(desc)), "steam_batch_info")
        self.launches_per_advance = launches.value
        self.descriptor_bytes = desc.value

    def __del__(self):
        try:
            if self.handle:
                _lib.load().steam_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    @property
    def voxels(self):
        return int(sum(inst.flux.numel() for inst in self.instances))

    def set_regions(self, regions):
        """New noise regions (keys) for the next advance, one per instance."""
        lib = _lib.load()
        for i, (inst, r) in enumerate(zip(self.instances, regions)):
            p = inst.native_noise(r)
            _lib.check(lib.steam_batch_set_noise(self.handle, i, ctypes.byref(p)),
                       "steam_batch_set_noise")

    def set_crops(self, crops):
        """Store only a box of each instance's advanced flux for the next
        advances: ``crops[i]`` is ``(out, origin)``, ``out`` a contiguous
        float32 device tensor of the box's shape and ``origin`` its corner in
        the instance's grid (the grid itself is then not written), or None
        for the whole grid.  Only for apply-mode instances with a unit
        entering flux, no S_k and a register z plan (served tiles); others
        raise ValueError."""
        lib = _lib.load()
        for i, c in enumerate(crops):
            if c is None:
                _lib.check(lib.steam_batch_set_crop(self.handle, i, None, None, None),
                           "steam_batch_set_crop")
                continue
            out, origin = c
            if (out.dtype != torch.float32 or not out.is_cuda or not out.is_contiguous()
                    or out.dim() != 3):
                raise ValueError("a crop is a contiguous float32 (x, y, z) device tensor")
            o = (ctypes.c_int64 * 3)(*[int(v) for v in origin])
            n = (ctypes.c_int64 * 3)(*[int(v) for v in out.shape])
            _lib.check(lib.steam_batch_set_crop(self.handle, i, _lib.ptr(out), o, n),
                       "steam_batch_set_crop")

    def advance(self, stream=None):
        """Advance every instance by one class on ``stream`` (default:
        current); asynchronous."""
        _lib.check(_lib.load().steam_batch_advance(self.handle,
                                                   _lib.stream_handle(stream)),
                   "steam_batch_advance")

    def advance_after_noise(self, stream=None):
        """``advance`` without its descriptor upload: the descriptors are the
        ones this batch's last ``generate_noise`` uploaded (every instance
        reads a noise buffer); asynchronous."""
        _lib.check(_lib.load().steam_batch_advance_after_noise(
            self.handle, _lib.stream_handle(stream)), "steam_batch_advance_after_noise")

    def generate_noise(self, stream=None):
        """Write every instance's noise for its next advance into its noise
        buffer (batches created with ``noise_buffers``); asynchronous."""
        _lib.check(_lib.load().steam_batch_noise(self.handle, _lib.stream_handle(stream)),
                   "steam_batch_noise")

    def mark_replay(self, stream=None):
        _lib.check(_lib.load().steam_batch_mark_replay(self.handle,
                                                       _lib.stream_handle(stream)),
                   "steam_batch_mark_replay")

    def capture(self):
        """One advance captured into a hipGraph (torch CUDAGraph), returned as
        a ``CapturedAdvance`` whose ``replay`` marks each replay for
        ``set_regions``."""
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.advance()  # warm the launch path outside the capture
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        with torch.cuda.graph(g):
            self.advance()
        return CapturedAdvance(self, g)

    def ledgers(self):
        """Each instance's ledger of the last advance (synchronises)."""
        lib = _lib.load()
        out = []
        for op, ws in zip(self.ops, self.workspaces):
            L = _lib.FluxLedger()
            _lib.check(lib.steam_flux_ledger_read(op.plan, ws.data_ptr(),
                                                  ctypes.byref(L),
                                                  _lib.stream_handle()),
                       "steam_flux_ledger_read")
            out.append(FluxAdvanceLedger.from_native(L))
        return out


class CapturedAdvance:
    """A captured batched advance.  A replay's descriptor upload reads the
    batch's pinned staging copy when the replay runs, so each ``replay``
    marks its end on the stream, and the batch's next ``set_regions`` waits
    for that mark before it rewrites the copy (no caller synchronisation
    needed between a replay and new keys)."""

    def __init__(self, batch, graph):
        self.batch = batch
        self.graph = graph

    def replay(self):
        self.graph.replay()
        _lib.check(_lib.load().steam_batch_mark_replay(self.batch.handle,
                                                       _lib.stream_handle()),
                   "steam_batch_mark_replay")


class PipelinedBatch:
    """The batched advance with each step's noise generated during the step
    before.  Two batches over the same instances (flux and kernel spectra
    shared) read precomputed noise from two buffer sets: step i advances
    with batch i % 2 on the current stream while, on a side stream, batch
    (i + 1) % 2 writes the noise of step i + 1 (VALU-bound work beside
    memory-bound transforms).  Each step is one fork and join, so a step
    captures into one graph (``capture``: one graph per parity, replayed in
    turn).  Results equal the plain batched advance with the same keys.

    A batch's descriptors go up once per use, with its noise: step i's
    advance runs on what batch i % 2's noise launch of step i - 1 uploaded.
    So a replay of step i reads only batch (i + 1) % 2's staging copy, and
    the keys of step i + 2 may be written as soon as step i's replay has
    run, while step i + 1 is still queued: the host never drains the device
    between steps.

    ``set_regions(regions)`` gives the keys of the next step whose noise is
    not yet generated (step i + 1 while step i is queued)."""

    def __init__(self, instances, autotune=True):
        import copy
        self.instances = list(instances)
        self.buffers = [[torch.empty_like(inst.flux) for inst in self.instances]
                        for _ in range(2)]
        # every step leaves its volume-mean rescale pending in the shared
        # workspace and the next step applies it on read (no separate rescale
        # pass); the first step has nothing pending to read
        chained, first = [], []
        for inst in self.instances:
            c = copy.copy(inst)
            c.defer_rescale, c.chain_pending = True, True
            chained.append(c)
            f = copy.copy(inst)
            f.defer_rescale, f.chain_pending = True, False
            first.append(f)
        b0 = BatchedAdvance(chained, autotune=autotune, noise_buffers=self.buffers[0])
        b1 = BatchedAdvance(chained, autotune=autotune, noise_buffers=self.buffers[1],
                            workspaces=b0.workspaces)
        self.first = BatchedAdvance(first, autotune=autotune, noise_buffers=self.buffers[0],
                                    workspaces=b0.workspaces)
        self.batches = [b0, b1]
        self.side = torch.cuda.Stream()
        self.step = 0
        self.graphs = None
        self.first.generate_noise()   # step 0's noise

    @property
    def voxels(self):
        return self.batches[0].voxels

    @property
    def launches_per_step(self):
        """The advance's launches (its descriptors are already on the device)
        plus the next step's noise (its descriptor upload and one noise
        launch: every instance shares the noise kernel)."""
        return self.batches[0].launches_per_advance - 1 + 2

    def set_regions(self, regions):
        """Keys of step ``self.step + 1`` (the next noise to be generated)."""
        self.batches[(self.step + 1) % 2].set_regions(regions)

    def _work(self, b, first=False):
        main = torch.cuda.current_stream()
        self.side.wait_stream(main)          # buffers 1 - b were last read by step i - 1
        with torch.cuda.stream(self.side):
            self.batches[1 - b].generate_noise(self.side)
        if first:
            self.first.advance(main)
        else:
            self.batches[b].advance_after_noise(main)
        main.wait_stream(self.side)

    def advance(self):
        """One step (eager)."""
        self._work(self.step % 2, first=self.step == 0)
        self.step += 1

    def finish(self):
        """Apply the last step's pending rescale to every flux (synchronises)."""
        for inst, led in zip(self.instances, self.ledgers()):
            r = led.rescale
            if r is None:
                inst.flux.fill_(np.float32(led.entering.mean))
            else:
                inst.flux.mul_(float(r))

    def capture(self):
        """Capture the two step graphs (parity 0 and 1), running nothing;
        ``replay`` then runs the next step from them.  At least one eager
        step must have run first (every kernel of the path loaded)."""
        if self.step < 1:
            raise ValueError("PipelinedBatch.capture: run one eager step first")
        torch.cuda.synchronize()
        self.graphs = [None, None]
        for b in (0, 1):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._work(b)
            self.graphs[b] = g
        return self.graphs

    def replay(self):
        self.graphs[self.step % 2].replay()
        # the replay's only descriptor upload read the staging copy of the
        # batch whose noise it generated
        self.batches[(self.step + 1) % 2].mark_replay()
        self.step += 1

    def ledgers(self):
        """Each instance's ledger of the last step (synchronises)."""
        return self.batches[0].ledgers()   # the workspaces are shared


def draw_batch_config(n=48, seed=48):
    """The batched workload's instance list: grid uniformly from
    {32^3, 64^3, 96x96x48, 128^3, 192x192x96, 256^3}, alpha ~ U[1.5, 1.9],
    C1 ~ U[0.05, 0.15] (c = C1^(1/alpha)), H ~ U[0.2, 0.5], seeds 0..n-1,
    drawn in that order from default_rng(seed)."""
    grids = [(32, 32, 32), (64, 64, 64), (96, 96, 48), (128, 128, 128),
             (192, 192, 96), (256, 256, 256)]
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        g = grids[int(rng.integers(len(grids)))]
        alpha = float(rng.uniform(1.5, 1.9))
        c1 = float(rng.uniform(0.05, 0.15))
        h = float(rng.uniform(0.2, 0.5))
        out.append({"seed": i, "grid": g, "alpha": alpha, "C1": c1,
                    "c": c1 ** (1.0 / alpha), "H": h})
    return out
