#!/usr/bin/env python3
"""Field-generation hot-path benchmark on one MI355X.

Workload (the plain GPU configuration of the headline metric): one flux-cascade
class advance of a 512 x 512 x 256 float32 field per step — keyed Philox noise
and CMS Levy variates at every cell (s = 1), alpha = 1.8, C1 = 0.1
(c = C1^(1/alpha)), a 65^3 Mexican-hat turbulon kernel, the periodic-x/y
zero-padded-z FFT convolution, clip, and the volume-mean restore (deferred and
chained on the device, as the resident cascade runs it).  The flux is resident
in HBM when the timed region starts.

Prints one JSON line (one GPU):
  value         voxels/s of the resident advance
  roofline      dominant kernel's algorithmic bytes / its average duration
                (HIP events on its stream, live over the timed region)
  cpu_baseline  the oracle's float32 restatement of the same class advance
                (C Philox/CMS on all host cores + torch-CPU FFT overlap-add,
                the reference's algorithm) on a bounded sub-box
  plain         the NumPy drop-in (_advance_flux with host arrays: PCIe in and
                out per call) on the same workload
"""

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("voxels/s (FP32 batched/served vs plain) at ≤1e-5 rel-L2; "
          "p50/p99 tile latency")
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-sample", type=int, nargs=3, default=None,
                    help="host sub-box for the CPU baseline (default: the grid)")
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched form (48 realisations)")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the served form (tile stream)")
    ap.add_argument("--serve-rate", type=float, default=1000.0,
                    help="tile requests per second of the served form")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-sweep", type=float, nargs="*", default=[500.0, 2000.0, 4000.0, 5000.0],
                    help="also serve the trace at these rates (req/s) and report "
                         "throughput and p50/p99 per rate")
    ap.add_argument("--serve-world", type=int, nargs=3, default=[2048, 2048, 256])
    ap.add_argument("--no-full-simulate", action="store_true",
                    help="skip the full-simulate form (flux + h + qt, all classes)")
    ap.add_argument("--full-simulate-only", action="store_true",
                    help="only run the full-simulate form; prints one JSON line")
    ap.add_argument("--production-square", action="store_true",
                    help="only time the scalar-path operations at 2048x2048x115; "
                         "prints one JSON line")
    ap.add_argument("--precision-form", action="store_true",
                    help="only run the precision form: one class advance of a "
                         "1024x1024x256 float32 field (alpha 1.6, C1 0.12, 129^3 "
                         "kernel) timed on the GPU and checked against the FP64 "
                         "oracle; prints one JSON line")
    ap.add_argument("--streamed-form", action="store_true",
                    help="only run the streamed form: the 2048x128x115 root cascade "
                         "resident and tile by tile; prints one JSON line")
    ap.add_argument("--refine-form", action="store_true",
                    help="only run the refine form: a nest of the 2048x128x115 root and a "
                         "nest of that nest; prints one JSON line")
    ap.add_argument("--pipelined-only", action="store_true",
                    help="only run the reported (pipelined) timed region of the class "
                         "advance, for profiling it alone; prints one JSON line")
    ap.add_argument("--batched-only", action="store_true",
                    help="only run the batched form (48 realisations); prints one "
                         "JSON line")
    ap.add_argument("--served-only", action="store_true",
                    help="only run the served form (tile stream at --serve-rate and "
                         "the --serve-sweep rates); prints one JSON line")
    ap.add_argument("--noise-ahead", type=int, default=2,
                    help="classes whose noise the pipelined form generates ahead")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy three times (the "
                         "reference kernel for scaling rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE into bytes)")
    return ap.parse_args()


def copy_calibration():
    import torch
    n = (1 << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device='cuda')
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4,
                      "launches": 3}))


def valu_ceiling(reps=3):
    """Measured compute-only VALU ceiling: wave-level fma instructions issued
    per second by the library's probe (8 independent fma chains per lane, 8
    waves per CU), best of ``reps``."""
    import torch
    from steam_amd import _lib
    lib = _lib.load()
    wgs, iters = 2048, 20000
    out = torch.empty(wgs * 256, dtype=torch.float32, device='cuda')
    n = ctypes.c_int64()
    best = None
    for _ in range(reps + 1):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        _lib.check(lib.steam_valu_probe(wgs, iters, _lib.ptr(out), ctypes.byref(n),
                                        _lib.stream_handle()), "steam_valu_probe")
        b.record()
        b.synchronize()
        rate = n.value / (a.elapsed_time(b) * 1e-3)
        best = rate if best is None else max(best, rate)
    return best


def read_valu_per_wave(path, kernel_prefix):
    """VALU instructions per wave and waves per dispatch of a kernel, from the
    committed SQ-counter summary."""
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        for line in fh:
            if line.startswith(kernel_prefix):
                f = line[74:].split()
                return float(f[4]), float(f[3])
    return None, None


def read_traffic(path):
    """Per-family HBM bytes per launch from the committed rocprofv3 summary
    (FETCH_SIZE / WRITE_SIZE scaled by the copy calibration)."""
    out = {}
    if not os.path.exists(path):
        return out
    with open(path) as fh:
        for line in fh:
            if line.startswith("#") or not line.strip():
                continue
            f = line.split()
            out[f[0]] = int(float(f[-3]))
    return out


def measure_copy_gbs(nbytes=1 << 30, reps=5):
    """Read+write bandwidth of a plain device copy (the practical HBM ceiling
    the kernels are compared with beside the 8 TB/s specification)."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device='cuda')
    dst = torch.empty_like(src)
    dst.copy_(src)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        dst.copy_(src)
    stop.record()
    stop.synchronize()
    ms = start.elapsed_time(stop) / reps
    del src, dst
    return 2 * nbytes / (ms * 1e-3) / 1e9


def family_bytes(info, shape, kernel_bytes):
    """Algorithmic HBM bytes per launch of each kernel family (fp32 state,
    complex64 spectra): the bytes the pass must move at minimum.
    ``kernel_bytes``: the kernel spectrum the x pass reads (complex64 per
    bin, or float32 per bin x <= nx/2 for a mirror-symmetric kernel)."""
    nx, ny, nz = shape
    N = nx * ny * nz
    E = info.spectral_elems
    return {
        "noise": 4 * N,                 # write the noise field
        "zfwd": 8 * N + 8 * E,          # read flux + noise, write planes
        "ypass": 16 * E,                # read + write planes
        "xpass": 16 * E + kernel_bytes,  # read + write planes, read kernel spectrum
        "zinv": 8 * E + 8 * N,          # read planes, read + write flux
    }


def kernel_spectrum_bytes(op, spec, shape):
    """Bytes of kernel spectrum one x pass reads, from the spectrum's format."""
    nx, ny, _ = shape
    nzc = op.n_fft // 2 + 1
    if op.spectrum_format(spec) == 'real_even':
        return 4 * nzc * (nx // 2 + 1) * ny
    return 8 * nzc * nx * ny


def cpu_baseline(sample, c, alpha, kernel_np, seed_key, repeats=3):
    """Oracle float32 class advance (reference algorithm) on host cores,
    best of ``repeats`` after one warm-up."""
    import numpy as np
    import torch
    from oracle import model as M
    cores = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        cores = min(cores, int(omp))
    torch.set_num_threads(cores)
    shape = tuple(sample)
    flux = np.ones(shape, np.float32)
    small = np.ones((32, 32, 32), np.float32)
    M.advance_flux(small, seed_key, (0, 0, 0), small.shape, kernel_np, c, 1,
                   (1, 1, 1), alpha=alpha, dtype=np.float32,
                   convolve=M.conv_oa_torch_f32)
    dt = float("inf")
    for _ in range(max(1, repeats)):
        t0 = time.perf_counter()
        M.advance_flux(flux, seed_key, (0, 0, 0), shape, kernel_np, c, 1,
                       (1, 1, 1), alpha=alpha, dtype=np.float32,
                       convolve=M.conv_oa_torch_f32)
        dt = min(dt, time.perf_counter() - t0)
    vox = int(np.prod(shape))
    return {"value": vox / dt, "unit": "voxels/s", "cores": cores,
            "kind": "port",
            "sample": (f"one class advance of a {shape[0]}x{shape[1]}x{shape[2]} "
                       f"float32 grid of the same workload (same kernel, alpha, "
                       f"c), best of {max(1, repeats)}: {dt:.2f} s")}


def batched_form(steps, warmup):
    """48 realisations of uneven size, alpha and noise scale advanced by one
    class per step: graph-captured batched advance vs the plain drop-in
    (one host-array _advance_flux call per realisation) and vs the resident
    single-instance advance called once per realisation."""
    import numpy as np
    import torch
    from steam_amd import simulate as S
    from steam_amd.batch import BatchedAdvance, Instance, PipelinedBatch, draw_batch_config

    cfg = draw_batch_config(48, seed=48)
    kern = S._turbulon_envelope(1, 0.5, 0.5, 0.5, as_tensor=True)
    insts = []
    for d in cfg:
        ss = np.random.SeedSequence(d["seed"])
        region = S.NoiseRegion(S.class_key(ss.spawn(1)[0]), (0, 0, 0), d["grid"])
        flux = torch.ones(d["grid"], dtype=torch.float32, device='cuda')
        insts.append(Instance(flux, kern, region, d["c"], alpha=d["alpha"]))
    t_setup = time.perf_counter()
    batch = BatchedAdvance(insts, autotune=True)
    setup_s = time.perf_counter() - t_setup
    graph = batch.capture()
    for _ in range(warmup):
        graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        graph.replay()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    # eager batched advance with fresh keys every step
    regions = [[S.NoiseRegion(S.class_key(np.random.SeedSequence(1000 * k + i)),
                              (0, 0, 0), d["grid"]) for i, d in enumerate(cfg)]
               for k in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        batch.set_regions(regions[k])
        batch.advance()
    torch.cuda.synchronize()
    ms_eager = (time.perf_counter() - t0) * 1e3 / steps
    vox = batch.voxels
    # resident single-instance advances, one call per realisation
    for i, inst in enumerate(insts):
        S.advance_flux_device(inst.flux, inst.region, kern, inst.flux_noise_scale, 1,
                              (1, 1, 1), alpha=inst.alpha, op=batch.ops[i],
                              spec=batch.specs[i], sync_ledger=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, inst in enumerate(insts):
        S.advance_flux_device(inst.flux, inst.region, kern, inst.flux_noise_scale, 1,
                              (1, 1, 1), alpha=inst.alpha, op=batch.ops[i],
                              spec=batch.specs[i], sync_ledger=False)
    torch.cuda.synchronize()
    single_s = time.perf_counter() - t0
    # plain drop-in: NumPy arrays through the reference's call, one per realisation
    kern_np = kern.cpu().numpy()
    hosts = [np.ones(d["grid"], np.float32) for d in cfg]
    S._advance_flux(hosts[0], insts[0].region, kern_np, insts[0].flux_noise_scale, 1,
                    (1, 1, 1), alpha=insts[0].alpha)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for h, inst in zip(hosts, insts):
        S._advance_flux(h, inst.region, kern_np, inst.flux_noise_scale, 1, (1, 1, 1),
                        alpha=inst.alpha)
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    # pipelined: each step's noise generated during the step before, on a
    # side stream (two graphs, one per buffer parity), fresh keys every step
    pipe = PipelinedBatch(insts, autotune=True)
    pregions = [[S.NoiseRegion(S.class_key(np.random.SeedSequence(5000 * k + i)),
                               (0, 0, 0), d["grid"]) for i, d in enumerate(cfg)]
                for k in range(warmup + steps + 2)]
    pipe.set_regions(pregions[0])
    pipe.advance()
    pipe.capture()
    for k in range(warmup):
        pipe.set_regions(pregions[k + 1])
        pipe.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        pipe.set_regions(pregions[warmup + k + 1])
        pipe.replay()
    torch.cuda.synchronize()
    ms_pipe = (time.perf_counter() - t0) * 1e3 / steps
    value = vox / (ms_pipe * 1e-3)
    graph_value = vox / (ms * 1e-3)
    e2e = 16 * vox / (ms_pipe * 1e-3) / 1e9
    return {
        "value": value, "unit": "voxels/s", "instances": len(insts), "voxels": vox,
        "ms_per_step": ms_pipe, "graph": True,
        "form": "pipelined: each step's noise generated beside the step before "
                "(two graphs), fresh keys every step",
        "launches_per_step": pipe.launches_per_step,
        "bytes_per_step": 16 * vox,
        "roofline_end_to_end": {"bytes_per_voxel": 16, "GBs": round(e2e, 1),
                                "frac": round(e2e / HBM_PEAK_GBS, 4)},
        "unpipelined_graph_ms_per_step": ms,
        "unpipelined_graph_voxels_per_s": graph_value,
        "unpipelined_launches_per_step": batch.launches_per_advance,
        "eager_new_keys_ms_per_step": ms_eager,
        "single_resident_voxels_per_s": vox / single_s,
        "vs_single_resident": value / (vox / single_s),
        "plain_dropin_voxels_per_s": vox / plain_s,
        "vs_plain_dropin": value / (vox / plain_s),
        "vs_plain_dropin_note": "the plain drop-in takes NumPy arrays and copies each "
                                "realisation host->device->host per call (PCIe-inclusive)",
        "n_fft_autotuned": sorted(set(batch.n_fft)), "setup_s": setup_s,
        "kernel": list(kern.shape),
        "workload": "48 realisations, grids from {32^3, 64^3, 96x96x48, 128^3, "
                    "192x192x96, 256^3}, alpha~U[1.5,1.9], C1~U[0.05,0.15], "
                    "default_rng(48)",
    }


def served_form(world, n_req, rate, sweep=None):
    """World-keyed tile stream: tiled world measure, then Poisson requests
    served with continuous batching; voxels/s and p50/p99 latency."""
    import numpy as np
    import torch
    from steam_amd.serve import TileServer, WorldClass, make_trace

    w = WorldClass(tuple(world), seed=0)
    t0 = time.perf_counter()
    w.measure(tile=(512, 512))
    torch.cuda.synchronize()
    measure_s = time.perf_counter() - t0
    warm = make_trace(16, rate, w.world_shape, seed=99)
    TileServer(w, cache_bytes=0).serve(warm)
    trace = make_trace(n_req, rate, w.world_shape, seed=0)
    srv = TileServer(w, budget_bytes=16 << 30, streams=4, cache_bytes=4 << 30)
    lat, wall, vox, _ = srv.serve(trace)
    curve = []
    for r in sweep or []:
        tr = make_trace(n_req, r, w.world_shape, seed=0)
        sv = TileServer(w, budget_bytes=16 << 30, streams=4, cache_bytes=4 << 30)
        la, wa, vo, _ = sv.serve(tr)
        curve.append({"rate_per_s": r, "voxels_per_s": vo / wa,
                      "p50_ms": float(np.percentile(la, 50) * 1e3),
                      "p99_ms": float(np.percentile(la, 99) * 1e3)})
    return {
        "sweep": curve,
        "value": vox / wall, "unit": "voxels/s", "requests": n_req,
        "rate_per_s": rate, "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3), "wall_s": wall,
        "world": list(w.world_shape), "kernel": list(w.kernel.shape),
        "measure_pass_s": measure_s, "cache_hits": srv.stats["cache_hits"],
        "max_inflight": srv.stats["max_inflight"],
        "groups": srv.stats["groups"], "max_group": srv.stats["max_group"],
        "mean_group": srv.stats["batched_requests"] / max(1, srv.stats["groups"]),
        "tiles": "{64^3, 128^3, 256^3, 384x384x192}, 70% periodic neighbours",
    }


class NoiseAhead:
    """The pipelined form's noise: class i's field is generated ``ahead``
    classes in advance on a side stream into a ring of ``ahead + 1`` buffers
    (the noise is flux-independent and VALU-bound), while the advances run
    their memory-bound FFT passes on the current stream; the side stream
    waits only for the advance that last read the buffer it refills."""

    def __init__(self, op, regions, c, alpha, shape, ahead):
        import torch
        from steam_amd import simulate as S
        self._gen_args = (S, op, regions, c, alpha)
        self.ahead, self.nbuf = ahead, ahead + 1
        self.main = torch.cuda.current_stream()
        self.side = torch.cuda.Stream()
        self.bufs = [torch.empty(shape, dtype=torch.float32, device='cuda')
                     for _ in range(self.nbuf)]
        self.ev_noise = [torch.cuda.Event() for _ in range(self.nbuf)]
        self.ev_adv = [torch.cuda.Event() for _ in range(self.nbuf)]
        self.queued = -1

    def _gen(self, i):
        S, op, regions, c, alpha = self._gen_args
        b = i % self.nbuf
        if i >= self.nbuf:
            self.side.wait_event(self.ev_adv[b])
        S.generate_noise_device(op, regions[i], c, 1, (1, 1, 1), self.bufs[b],
                                alpha=alpha, stream=self.side)
        self.ev_noise[b].record(self.side)
        self.queued = i

    def step(self, i, advance):
        """Queue class i: its noise (and the next ones' up to ``ahead``) on
        the side stream, then ``advance(noise_t)`` on the current stream."""
        b = i % self.nbuf
        while self.queued < i + self.ahead:
            self._gen(self.queued + 1)
        self.main.wait_event(self.ev_noise[b])
        advance(self.bufs[b])
        self.ev_adv[b].record(self.main)


def precision_form(steps=10):
    """The precision workload: float32 FFTs, spectral multiply and cascade
    arithmetic; float64 only in the kernel-correction sums and the volume
    sums / rescale (all on the GPU); checked against the FP64 oracle on the
    same keyed uniforms (flux and S_k: rel-L2 and max-abs), and the
    integrated field (one class's scalar increment) likewise."""
    import numpy as np
    import torch
    from oracle import model as M
    from steam_amd import simulate as S
    from steam_amd.utils import conv_operator

    def say(msg):
        print(f"[precision form] {msg}", file=sys.stderr, flush=True)

    shape = (1024, 1024, 256)
    alpha, c1 = 1.6, 0.12
    c = c1 ** (1.0 / alpha)
    key = S.class_key(np.random.SeedSequence(0).spawn(1)[0])
    region = S.NoiseRegion(key, (0, 0, 0), shape)
    kern = S._turbulon_envelope(1, 0.5, 0.5, 0.5, support_factor=32, as_tensor=True)
    rng = np.random.default_rng(1)
    flux0 = np.abs(1 + 0.3 * rng.standard_normal(shape, dtype=np.float32))
    op = conv_operator(shape, tuple(kern.shape), precision='kahan')
    spec = op.spectrum(kern)
    say(f"kernel {tuple(kern.shape)}, n_fft {op.n_fft}, blocks {op.n_blocks}")
    ft = torch.from_numpy(flux0).cuda()
    sk = torch.empty_like(ft)
    S.advance_flux_device(ft, region, kern, c, 1, (1, 1, 1), alpha=alpha, op=op,
                          spec=spec, sk_t=sk, precision='kahan')
    f32 = ft.cpu().numpy()
    s32 = sk.cpu().numpy()
    # timing: deferred, chained advances of the resident field
    work = torch.from_numpy(flux0).cuda()
    for i in range(2):
        S.advance_flux_device(work, region, kern, c, 1, (1, 1, 1), alpha=alpha, op=op,
                              spec=spec, defer_rescale=True, sync_ledger=False,
                              chain_pending=i > 0, precision='kahan')
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        S.advance_flux_device(work, region, kern, c, 1, (1, 1, 1), alpha=alpha, op=op,
                              spec=spec, defer_rescale=True, sync_ledger=False,
                              chain_pending=True, precision='kahan')
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    # the same chained advances with each class's noise generated two
    # classes ahead beside the advance before (fresh keys per class)
    seeds = np.random.SeedSequence(7).spawn(2 * steps + 8)
    regions = [S.NoiseRegion(S.class_key(sq), (0, 0, 0), shape) for sq in seeds]
    ahead_noise = NoiseAhead(op, regions, c, alpha, shape, 2)

    def piped(i):
        ahead_noise.step(i, lambda noise_t: S.advance_flux_device(
            work, regions[i], kern, c, 1, (1, 1, 1), alpha=alpha, op=op, spec=spec,
            defer_rescale=True, sync_ledger=False, chain_pending=True, precision='kahan',
            noise_t=noise_t))
    for i in range(2):
        piped(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(2, 2 + steps):
        piped(i)
    torch.cuda.synchronize()
    ms_piped = (time.perf_counter() - t0) * 1e3 / steps
    del ahead_noise
    say(f"GPU {ms:.3f} ms per advance, {ms_piped:.3f} pipelined; running the FP64 oracle")
    kern64 = M.turbulon_envelope(1, 0.5, 0.5, 0.5, support_factor=32, dtype=np.float64)
    t0 = time.perf_counter()
    f64, s64, led, _ = M.advance_flux(flux0, key, (0, 0, 0), shape, kern64, c, 1,
                                      (1, 1, 1), alpha=alpha, dtype=np.float64)
    oracle_s = time.perf_counter() - t0
    say(f"oracle {oracle_s:.1f} s")

    def err(a, b):
        d = a.astype(np.float64) - b
        return (float(np.sqrt((d * d).sum() / (b * b).sum())), float(np.abs(d).max()))

    fl2, fmax = err(f32, f64)
    sl2, smax = err(s32, s64)
    vox = int(np.prod(shape))

    # the integrated field: one class's scalar increment T * (C_k P / <|P|>)
    # with P the joint pattern of a running field on this class's S_k; the
    # GPU in float32 from its own S_k, the oracle in float64 from its own
    from oracle import scalars as OS
    from steam_amd import scalars as SC
    nz = shape[2]
    zc = np.arange(nz) * 0.5
    pert = (50.0 * np.random.default_rng(2).standard_normal(shape, dtype=np.float32))
    pert = pert.astype(np.float32)
    mean = np.linspace(340e3, 320e3, nz).astype(np.float32)
    c_k = np.full(nz, 500.0, dtype=np.float32)
    b_z = np.full(nz, 3000.0, dtype=np.float32)
    aspect = np.full(nz, 0.3, dtype=np.float32)
    lo, hi = 315 * 1004, 355 * 1004
    pt = torch.from_numpy(pert).cuda()
    inc = torch.empty_like(pt)

    def scalar_increment():
        W = SC.joint_pattern(pt, mean, sk, b_z, aspect, zc, 0.5, 0.5, lo, hi, precise=True)
        led = SC.product_norm(W)
        SC.normalize_pattern(W, led.level_mean, c_k)
        op.execute(W, inc, spec)

    scalar_increment()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        scalar_increment()
    torch.cuda.synchronize()
    inc_ms = (time.perf_counter() - t0) * 1e3 / 3
    i32 = inc.cpu().numpy()
    say(f"GPU scalar increment {inc_ms:.3f} ms; running the FP64 oracle for it")
    W64 = OS.joint_pattern(pert, mean, s64, b_z, aspect, zc, 0.5, 0.5, lo, hi,
                           dtype=np.float64)
    _, _, lm64 = OS.product_norm(W64, dtype=np.float64)
    W64 = OS.normalize(W64, lm64, c_k, dtype=np.float64)
    say("oracle pattern done; convolving it in float64")
    i64 = M.conv_fft(W64, kern64)
    del W64
    il2, imax = err(i32, i64)
    return {"metric": "precision form: float32 class advance vs FP64 oracle",
            "workload": "flux_advance_class_1024x1024x256", "grid": list(shape),
            "kernel": list(kern.shape), "alpha": alpha, "C1": c1, "c": c,
            "n_fft": op.n_fft, "n_blocks": op.n_blocks,
            "ms_per_step": ms, "voxels_per_s": vox / (ms * 1e-3),
            "ms_per_step_pipelined": ms_piped,
            "voxels_per_s_pipelined": vox / (ms_piped * 1e-3),
            "pipelined": "each class's noise generated two classes ahead on a side stream "
                         "(ms_per_step: the noise generated inside each advance)",
            "flux_rel_l2": fl2, "flux_max_abs": fmax,
            "sk_rel_l2": sl2, "sk_max_abs": smax,
            "flux_max_abs_relative_to_max": fmax / float(np.abs(f64).max()),
            "integrated_rel_l2": il2, "integrated_max_abs": imax,
            "integrated_max_abs_relative_to_max": imax / float(np.abs(i64).max()),
            "integrated_ms": inc_ms,
            "integrated_field": "one class's scalar increment T * (C_k P / <|P|>_level), "
                                "P = W S_k g of a running field 320-340 kJ/kg + 50 N(0,1) "
                                "(default_rng(2)), C_k 500, b 3000, aspect 0.3",
            "oracle_s": oracle_s,
            "data": "entering flux |1 + 0.3 N(0,1)| from default_rng(1), key "
                    "SeedSequence(0).spawn(1)[0]"}


# The reference's published cost of a full simulate() (flux + h + qt, all
# classes) on a 16-core AMD Ryzen 9950X, CPU only (paper/supplement.tex:713-716).
PUBLISHED_FULL_SIMULATE_S = {(1024, 64, 78): 1.7, (2048, 128, 115): 3.4}


def full_simulate_form(reps=3):
    """The whole root run on the GPU (set-up, cascade over every class
    carrying h, qt and flux, composition of the written fields) at the two
    grids of the reference's cost comparison: a 384 km x 6144 km channel at
    6 km (1024x64x78) and 3 km (2048x128x115) spacing, outer scale 384 km,
    spheroscale 10 m, 20 km deep.  The mean profiles are synthetic (the
    reference's were taken from CM1 output), so the grid ladder and the work
    match and the fields do not."""
    import numpy as np
    import torch
    from steam_amd import cascade as C

    z = np.arange(201) * 100.0
    hp = 340e3 - 20e3 * z / z.max()
    qp = 0.018 - 0.016 * z / z.max()
    out = []
    for nx, ny, dx in ((1024, 64, 6000.0), (2048, 128, 3000.0)):
        def run(seed):
            return C.simulate_fields(hp, qp, nx, ny, dx, dx, 384000.0, 10.0, 20000.0, 100.0,
                                     seed=seed, as_tensors=True)
        res = run(0)
        torch.cuda.synchronize()
        times = []
        for r in range(reps):
            t0 = time.perf_counter()
            res = run(r + 1)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        grids = res['grids']
        finest = tuple(int(v) for v in res['h'].shape)
        class_voxels = int(sum(int(a) * int(b) * int(c) for a, b, c in
                               zip(grids['nx'], grids['ny'], grids['nz'])))
        best = min(times)
        pub = PUBLISHED_FULL_SIMULATE_S.get(finest)
        out.append({"finest": list(finest), "classes": len(grids['k']),
                    "class_voxels": class_voxels, "best_ms": best * 1e3,
                    "median_ms": sorted(times)[len(times) // 2] * 1e3,
                    "finest_voxels_per_s": float(np.prod(finest)) / best,
                    "published_cpu_s": pub,
                    "speedup_vs_published_cpu": (pub / best) if pub else None})
    return {"workload": "full simulate (flux + h + qt), reference cost-comparison grids",
            "published_cpu": "16-core AMD Ryzen 9950X (paper/supplement.tex:713-716)",
            "data": "synthetic linear h/qt profiles, seeds 0..reps", "runs": out}


def streamed_form(reps=3, plans=((5, 2, 1), (4, 4, 2), (4, 8, 4))):
    """The whole root cascade (flux, h, qt and the deficits) resident and
    streamed tile by tile (streamed_cascade_loop: per tile the halo regrid,
    the flux advance with the raw-product pattern measure, the per-tile
    convolution of the normalised pattern; per-level passes on the whole
    fields) on the 2048x128x115 ladder of the cost-comparison grids.  Each
    plan is (horizon, tiles_x, tiles_y): the classes from the horizon on are
    tiled, the narrower ones before it run as one tile (plan_tiling's forced
    plan).  Reports the times, the tiling overhead, the streamed fields'
    rel-L2 against the resident ones, and the h differences in units of the
    float32 spacing of phi = perturbation + mean, the resolution h is carried
    at."""
    import numpy as np
    import torch
    from steam_amd.cascade import cascade_loop
    from steam_amd.params import root_setup
    from steam_amd.streaming import plan_tiling, streamed_cascade_loop

    z = np.arange(201) * 100.0
    hp = 340e3 - 20e3 * z / z.max()
    qp = 0.018 - 0.016 * z / z.max()
    st = root_setup(hp, qp, 2048, 128, 3000.0, 3000.0, 384000.0, 10.0, 20000.0, 100.0,
                    seed=0)
    hmin, hmax, qmin, qmax = st['bounds']
    args = (st['h_profile'], st['qt_profile'], st['z_profile'], st['grids'], st['C_h_k'],
            st['C_qt_k'], st['b_h_k'], st['b_qt_k'], st['aspect_k'], hmin, hmax, qmin, qmax,
            st['min_distance_to_ground'], st['sparsity_factors'], st['class_seeds'])
    kw = dict(comp_k=st['comp_k'], as_tensors=True)

    def timed(fn):
        res = fn()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            res = fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return res, best

    resident, t_res = timed(lambda: cascade_loop(*args, **kw))
    # the resolution the h state is carried at: the bounded add forms
    # phi = perturbation + mean in float32, so h moves in steps of ulp(phi)
    z_fin = np.asarray(resident[6]['z'], dtype=np.float32)
    h_mean = torch.from_numpy(np.interp(z_fin, st['z_profile'], st['h_profile'])
                              .astype(np.float32)).cuda()
    phi = (resident[0] + h_mean).abs()
    ulp_phi = (torch.nextafter(phi, torch.full_like(phi, float("inf"))) - phi).double()
    runs = []
    for force in plans:
        tiles = plan_tiling(st['grids'], (13, 13, 13), force=force)
        streamed, t = timed(lambda: streamed_cascade_loop(*args, tiles, **kw))
        rel = {}
        for name, k in (("h_perturbation", 0), ("qt_perturbation", 1), ("flux_state", 2)):
            a = streamed[k].double()
            b = resident[k].double()
            rel[name] = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
        steps = (streamed[0].double() - resident[0].double()).abs() / ulp_phi
        runs.append({"plan": {"horizon": force[0], "tiles": [force[1], force[2]]},
                     "best_ms": t * 1e3, "vs_resident": t / t_res,
                     "rel_l2_vs_resident": rel,
                     "h_difference_in_ulps_of_phi": {
                         "max": float(steps.max()),
                         "cells_above_1": float((steps > 1.0).double().mean()),
                         "cells_above_4": float((steps > 4.0).double().mean())}})
    g = st['grids']
    return {"workload": "root cascade 2048x128x115 (7 classes), resident and streamed",
            "class_voxels": int(sum(int(a) * int(b) * int(c)
                                    for a, b, c in zip(g['nx'], g['ny'], g['nz']))),
            "resident_ms": t_res * 1e3, "streamed": runs,
            "h_rms_over_ulp_phi": float(resident[0].double().pow(2).mean().sqrt()
                                        / ulp_phi.mean()),
            "data": "synthetic linear h/qt profiles, seed 0"}


def refine_form(reps=3):
    """refine at the size of the cost-comparison runs: the 2048x128x115 root
    (saved for refinement) refined over x in [512, 1536), y in [32, 96) to
    1.5 km (one class, the halo padding the nest to 2060x140 columns), and
    that nest refined again over its middle to 750 m; device-resident
    records and outputs (the nest also from the record's NumPy copy).  Checks the bounds and that a replay from the nest's ledger is
    bit-identical."""
    import numpy as np
    import torch
    from steam_amd import cascade as C
    from steam_amd.refine import refine_fields

    z = np.arange(201) * 100.0
    hp = 340e3 - 20e3 * z / z.max()
    qp = 0.018 - 0.016 * z / z.max()
    root = C.simulate_fields(hp, qp, 2048, 128, 3000.0, 3000.0, 384000.0, 10.0, 20000.0,
                             100.0, seed=0, save_for_refinement=True, as_tensors=True)
    parent = root['parent']
    # the same record as NumPy arrays (what a file read gives)
    host_parent = {k: (v.cpu().numpy() if isinstance(v, torch.Tensor) else v)
                   for k, v in parent.items()}
    host_parent['class_increments'] = [
        {k: (v.cpu().numpy() if isinstance(v, torch.Tensor) else v) for k, v in c.items()}
        for c in parent['class_increments']]

    def timed(fn):
        out = fn()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return out, best

    nest, t1 = timed(lambda: refine_fields(parent, 512, 1536, 32, 96, 1500.0, 1500.0,
                                           as_tensors=True))
    _, t1h = timed(lambda: refine_fields(host_parent, 512, 1536, 32, 96, 1500.0, 1500.0,
                                         as_tensors=True))
    replay = refine_fields(parent, 512, 1536, 32, 96, 1500.0, 1500.0, as_tensors=True,
                           ledger=nest['ledger'])
    identical = all(torch.equal(replay[k], nest[k]) for k in ('h', 'qt', 'flux'))
    hmin, hmax, qmin, qmax = parent['bounds']
    in_bounds = bool(nest['h'].min() >= hmin and nest['h'].max() <= hmax
                     and nest['qt'].min() >= qmin and nest['qt'].max() <= qmax
                     and nest['flux'].min() >= 0)
    rec = refine_fields(parent, 512, 1536, 32, 96, 1500.0, 1500.0, as_tensors=True,
                        save_for_refinement=True)['parent']
    nx = rec['nx']
    nn, t2 = timed(lambda: refine_fields(rec, nx // 4, nx // 4 + 512, 16, 112, 750.0, 750.0,
                                         as_tensors=True))
    return {"workload": "refine of the 2048x128x115 root and of its nest",
            "nest": {"region": "x [512, 1536), y [32, 96) of the root, to 1.5 km",
                     "shape": list(nest['h'].shape), "best_ms": t1 * 1e3,
                     "best_ms_from_host_record": t1h * 1e3,
                     "replay_bit_identical": identical, "in_bounds": in_bounds},
            "nest_of_nest": {"region": "x [nx/4, nx/4 + 512), y [16, 112) of the nest, "
                                       "to 750 m",
                             "shape": list(nn['h'].shape), "best_ms": t2 * 1e3},
            "data": "synthetic linear h/qt profiles, seed 0"}


# The reference's own quoted costs at its production square 2048x2048x115
# (an unnamed CUDA card; the bounded add and the projection include their
# host/device transfers): steam/utils.py:89-90, steam/simulate.py:2995, 3311.
PUBLISHED_PRODUCTION_S = {"convolve": 0.64, "bounded_add": 1.5, "projection_qt": 6.28}


def production_form(reps=3):
    """The scalar-path operations at the reference's production square
    2048x2048x115 (a 13^3 kernel at s = 1), timed on device-resident fields:
    CONVOLVE, the bounded amplitude-preserving add, the joint pattern with
    its product norm and normalisation, and the projection of a qt-like
    field whose levels mostly violate qt_min = 0.  The fields are synthetic
    (seeded normal draws), so the rescale loop and the violation pattern are
    those of this data, not of a cascade state."""
    import numpy as np
    import torch
    from steam_amd import cascade as C
    from steam_amd import scalars as SC
    from steam_amd import simulate as S
    from steam_amd.utils import convolve_fft_xy_oa_z

    shape = (2048, 2048, 115)
    gen = torch.Generator(device='cuda')
    gen.manual_seed(0)

    def rand(scale):
        return (scale * torch.randn(shape, device='cuda', generator=gen)).contiguous()

    def timed(fn, setup=None):
        args = setup() if setup else ()
        fn(*args)
        torch.cuda.synchronize()
        best = None
        for _ in range(reps):
            args = setup() if setup else ()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn(*args)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best

    field = rand(1.0)
    kern = S._turbulon_envelope(1, 0.5, 0.5, 0.5, as_tensor=True)
    conv_s = timed(lambda: convolve_fft_xy_oa_z(field, kern))
    del field
    mean = np.linspace(340e3, 320e3, 115).astype(np.float32)
    lo, hi = 315 * 1004, 355 * 1004
    pert0, inc0 = rand(300.0), rand(200.0)
    ba_s = timed(lambda p, d: SC.bounded_amplitude_add(p, mean, d, lo, hi, record_applied=True),
                 lambda: (pert0.clone(), inc0.clone()))
    sk = rand(1.0)
    z = np.arange(115) * 100.0
    b_z = np.full(115, 3000.0, np.float32)
    aspect = np.full(115, 0.3, np.float32)
    c_k = np.full(115, 500.0, np.float32)

    def pattern():
        W = SC.joint_pattern(pert0, mean, sk, b_z, aspect, z, 3000.0, 3000.0, lo, hi)
        SC.normalize_pattern(W, SC.product_norm(W).level_mean, c_k)

    pat_s = timed(pattern)
    del pert0, inc0, sk
    qmean = np.linspace(0.018, 0.002, 115).astype(np.float32)
    qt0 = rand(0.004)
    proj_s = timed(lambda q: C.project_onto_bounds(q, qmean, 0.0, 0.03), lambda: (qt0.clone(),))
    out = {"convolve": conv_s, "bounded_add": ba_s, "pattern_norm_normalise": pat_s,
           "projection_qt": proj_s}
    return {"workload": "scalar-path operations at 2048x2048x115 (13^3 kernel)",
            "best_ms": {k: v * 1e3 for k, v in out.items()},
            "reference_quoted_s": PUBLISHED_PRODUCTION_S,
            "ratio_vs_reference_quoted": {k: PUBLISHED_PRODUCTION_S[k] / out[k]
                                          for k in PUBLISHED_PRODUCTION_S},
            "data": "synthetic: h-like field 340-320 kJ/kg + 300 N(0,1), increment "
                    "200 N(0,1), qt-like 0.018-0.002 + 0.004 N(0,1), torch seed 0"}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.precision_form:
        from steam_amd import _lib
        _lib.require_gpu('cuda')
        print(json.dumps(precision_form()))
        return
    if args.production_square:
        from steam_amd import _lib
        _lib.require_gpu('cuda')
        print(json.dumps(production_form()))
        return
    if args.full_simulate_only:
        from steam_amd import _lib
        _lib.require_gpu('cuda')
        print(json.dumps(full_simulate_form()))
        return
    if args.streamed_form:
        from steam_amd import _lib
        _lib.require_gpu('cuda')
        print(json.dumps(streamed_form()))
        return
    if args.refine_form:
        from steam_amd import _lib
        _lib.require_gpu('cuda')
        print(json.dumps(refine_form()))
        return
    if args.batched_only:
        from steam_amd import _lib
        _lib.require_gpu('cuda')
        print(json.dumps(batched_form(args.steps, args.warmup)))
        return
    if args.served_only:
        from steam_amd import _lib
        _lib.require_gpu('cuda')
        print(json.dumps(served_form(args.serve_world, args.serve_requests,
                                     args.serve_rate, args.serve_sweep)))
        return
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (got --gpus {args.gpus}); the "
              f"distributed decomposition is designed, not built. Run with "
              f"--gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from steam_amd import _lib
    from steam_amd import simulate as S
    from steam_amd.utils import conv_operator

    _lib.require_gpu('cuda')
    lib = _lib.load()
    shape = (args.nx, args.ny, args.nz)
    alpha = 1.8
    c = 0.1 ** (1.0 / alpha)
    key = S.class_key(np.random.SeedSequence(0).spawn(1)[0])
    region = S.NoiseRegion(key, (0, 0, 0), shape)
    kern = S._turbulon_envelope(1, 0.5, 0.5, 0.5, support_factor=16,
                                as_tensor=True)
    flux = torch.ones(shape, dtype=torch.float32, device='cuda')
    op = conv_operator(shape, tuple(kern.shape))
    spec = op.spectrum(kern)
    info = _lib.ConvInfo()
    lib.steam_conv_plan_info(op.plan, ctypes.byref(info))

    total = args.warmup + args.steps
    seeds = np.random.SeedSequence(0).spawn(total + args.steps + args.noise_ahead + 4)
    regions = [S.NoiseRegion(S.class_key(sq), (0, 0, 0), shape) for sq in seeds]

    # serial form: each advance generates its own noise
    def serial_step(i):
        S.advance_flux_device(flux, regions[i], kern, c, 1, (1, 1, 1), alpha=alpha,
                              defer_rescale=True, sync_ledger=False, op=op,
                              spec=spec, chain_pending=i > 0)

    def read_profile():
        fam = {}
        for name, fid in _lib.KERNEL_FAMILIES.items():
            n = ctypes.c_int64()
            ms = ctypes.c_double()
            _lib.check(lib.steam_profile_read(fid, ctypes.byref(n), ctypes.byref(ms)))
            if n.value:
                fam[name] = {"launches": n.value, "avg_ms": ms.value / n.value,
                             "total_ms": ms.value}
        return fam

    # serial timed region: per-kernel HIP-event durations without overlap
    fam = {}
    serial_ms = None
    if not args.pipelined_only:
        for i in range(args.warmup):
            serial_step(i)
        torch.cuda.synchronize()
        _lib.check(lib.steam_profile_enable(1))
        t0 = time.perf_counter()
        for i in range(args.steps):
            serial_step(args.warmup + i)
        torch.cuda.synchronize()
        serial_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        fam = read_profile()
        _lib.check(lib.steam_profile_enable(0))

    # pipelined form: the noise fields of the next classes (flux-independent,
    # VALU-bound) are generated on a side stream, ``ahead`` classes in
    # advance into a ring of ahead + 1 buffers, while class i runs its
    # memory-bound FFT passes on the main stream; the side stream only waits
    # for the advance that last read the buffer it refills
    ahead_noise = NoiseAhead(op, regions, c, alpha, shape, max(1, args.noise_ahead))

    def piped_step(i):
        ahead_noise.step(i, lambda noise_t: S.advance_flux_device(
            flux, regions[i], kern, c, 1, (1, 1, 1), alpha=alpha, defer_rescale=True,
            sync_ledger=False, op=op, spec=spec, chain_pending=True, noise_t=noise_t))

    # pipelined timed region (the reported value), uninstrumented; every
    # pipelined advance chains the pending rescale the previous one left, so
    # the chain starts from one unchained advance when the serial region
    # did not run
    if args.pipelined_only:
        serial_step(0)
    for i in range(args.warmup):
        piped_step(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.warmup, total):
        piped_step(i)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    led, _ = S.read_device_ledger(shape, tuple(kern.shape))
    assert np.isfinite(led.realized.mean) and led.realized.mean > 0
    ms_step = (t1 - t0) * 1e3 / args.steps
    voxels = int(np.prod(shape))
    value = voxels / (ms_step * 1e-3)
    if args.pipelined_only:
        print(json.dumps({"form": "pipelined class advance only", "grid": list(shape),
                          "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": ms_step, "voxels_per_s": value}))
        return

    # the same pipelined region again with every launch bracketed by HIP
    # events on its own stream: the per-pass durations of the reported form
    # (the events add a little host work per launch, so the value above comes
    # from the uninstrumented run)
    _lib.check(lib.steam_profile_enable(1))
    t0 = time.perf_counter()
    for i in range(total, total + args.steps):
        piped_step(i)
    torch.cuda.synchronize()
    piped_instr_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    fam_piped = read_profile()
    _lib.check(lib.steam_profile_enable(0))

    kbytes = kernel_spectrum_bytes(op, spec, shape)
    fbytes = family_bytes(info, shape, kbytes)

    def per_pass(f):
        return {k: {"avg_ms": round(f[k]["avg_ms"], 5),
                    "ms_per_step": round(f[k]["total_ms"] / args.steps, 5),
                    "GBs": round(fbytes[k] / (f[k]["avg_ms"] * 1e-3) / 1e9, 1),
                    "frac": round(fbytes[k] / (f[k]["avg_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
                for k in fbytes if k in f and k != "noise"}
    # dominant memory-bound pass of the reported form: the most device time
    # per step on the main stream (the noise kernel runs beside the passes on
    # its own stream and is VALU-bound)
    dom = max((k for k in fam_piped if k in fbytes and k != "noise"),
              key=lambda k: fam_piped[k]["total_ms"])
    achieved = fbytes[dom] / (fam_piped[dom]["avg_ms"] * 1e-3) / 1e9
    traffic = None
    if shape == (512, 512, 256):
        traffic = read_traffic(os.path.join(ROOT, "profiles",
                                            "traffic_flux_advance_512x512x256.txt")).get(dom)
    copy_gbs = measure_copy_gbs()
    # end to end: 16 algorithmic bytes per voxel per class (read the entering
    # flux, write the advanced flux, the convolution's input and output)
    e2e_gbs = 16 * voxels / (ms_step * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes_per_launch": fbytes[dom],
                "avg_launch_ms": round(fam_piped[dom]["avg_ms"], 5),
                "timed_in": "pipelined region (HIP events on the launch stream)",
                "measured_copy_GBs": round(copy_gbs, 1),
                "frac_of_copy": round(achieved / copy_gbs, 4),
                "per_pass_pipelined": per_pass(fam_piped),
                "per_pass_serial": per_pass(fam),
                "end_to_end": {"bytes_per_voxel": 16, "GBs": round(e2e_gbs, 1),
                               "frac": round(e2e_gbs / HBM_PEAK_GBS, 4)},
                "instrumented_ms_per_step": round(piped_instr_ms, 4)}

    plain = None
    if not args.no_plain:
        host = np.ones(shape, np.float32)
        kern_np = kern.cpu().numpy()
        S._advance_flux(host, region, kern_np, c, 1, (1, 1, 1), alpha=alpha)
        torch.cuda.synchronize()
        reps = 3
        p0 = time.perf_counter()
        for _ in range(reps):
            S._advance_flux(host, region, kern_np, c, 1, (1, 1, 1), alpha=alpha)
        torch.cuda.synchronize()
        plain = voxels * reps / (time.perf_counter() - p0)

    batched = None if args.no_batched else batched_form(40, 4)
    served = None if args.no_serve else served_form(args.serve_world,
                                                     args.serve_requests,
                                                     args.serve_rate,
                                                     args.serve_sweep)

    full = None if args.no_full_simulate else full_simulate_form()

    # the VALU-bound noise kernel against the measured VALU ceiling: its
    # vector instructions per wave (committed SQ counters of the full-grid
    # launch of the serial region) over its measured serial duration
    noise_valu = None
    vpw, waves = read_valu_per_wave(
        os.path.join(ROOT, "profiles", "sq_counters_flux_advance_512x512x256.txt"),
        "void steam::noise_field_kernel<steam::SingleSrc, false")
    if vpw and fam.get("noise"):
        ceiling = valu_ceiling()
        achieved = vpw * waves / (fam["noise"]["avg_ms"] * 1e-3)
        noise_valu = {"bound": "valu", "valu_per_wave": vpw, "waves": waves,
                      "valu_per_centre": round(vpw * waves * 64 / voxels, 2),
                      "achieved_wave_instr_per_s": achieved,
                      "ceiling_wave_instr_per_s": ceiling,
                      "frac": round(achieved / ceiling, 4),
                      "note": "SQ_INSTS_VALU counts an instruction once whatever its rate; "
                              "the noise's 32x32->64 multiplies and transcendentals take "
                              "several issue slots each, so frac understates its VALU use"}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_sample or list(shape), c, alpha,
                           kern.cpu().numpy(), key, args.cpu_repeats)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (world-keyed Philox noise from seed 0, flux from 1)",
        "config": {"workload": "flux_advance_class_512x512x256",
                   "grid": list(shape), "kernel": list(kern.shape),
                   "sparsity": [1, 1, 1], "alpha": alpha, "C1": 0.1, "c": c,
                   "n_fft": info.n_fft, "n_blocks": info.n_blocks,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "noise_valu": noise_valu,
        "cpu_baseline": cpu,
        "plain_dropin_voxels_per_s": plain,
        "serial_resident_voxels_per_s": voxels / (serial_ms * 1e-3),
        "serial_resident_ms_per_step": serial_ms,
        "kernels": fam,
        "kernels_pipelined": fam_piped,
        "batched": batched,
        "served": served,
        "full_simulate": full,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
