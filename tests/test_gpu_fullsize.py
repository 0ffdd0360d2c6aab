"""GPU parity at the benchmark's full sizes: the 6000 x 4000 chain end to end, a batch spanning the
batched workload's frame sizes, and the recursive column passes at the heights the benchmarks run.

Each render records its grain draws (override / option "record_grain"), so the integer Poisson
counts are checked draw by draw:
  - where the GPU's Poisson mean equals the oracle's, the counts are identical;
  - every GPU count equals the oracle sampler's draw at the GPU's own mean (same pixel, stream and
    seed), so the compacted drain of the develop kernel is bit-exact wherever it runs;
  - the means themselves agree except where a recursive halation pass rounded 1 ulp apart from the
    oracle's serial recursion (bounded in count and size).
The RGBA16 output is compared with the oracle's own chain from the input (not the GPU's negative):
within 1 LSB except at pixels near a draw whose mean moved by such an ulp.
"""
import numpy as np
import pytest

import spektralab_engine_amd as spk
from conftest import GRAIN_HALATION, MIX3_OVERRIDES
from spektralab_engine_amd import synth

pytestmark = pytest.mark.gpu

FIRST_FILM_SEED = 0x4ABEA221  # splitmix64's first draw from the sequence's fixed start
STREAMS = np.array([c + 10 * l for c in range(3) for l in range(3)], np.uint32)  # draw 3 c + l


def ulps(a, b):
    ia = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def dilate(mask, r):
    """Boolean mask grown by r pixels in x and y (the grain blur's reach)."""
    out = mask.copy()
    h, w = mask.shape
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            ys, yd = (slice(max(dy, 0), h), slice(0, h - max(dy, 0))) if dy >= 0 else (slice(0, h + dy), slice(-dy, h))
            xs, xd = (slice(max(dx, 0), w), slice(0, w - max(dx, 0))) if dx >= 0 else (slice(0, w + dx), slice(-dx, w))
            out[yd, xd] |= mask[ys, xs]
    return out


def check_frame(oracle, name, plan, img, out, lam_g, cnt_g):
    """The draw-by-draw and end-to-end checks of one frame; returns a summary dict."""
    w, h = oracle.plan_size(plan)
    seed = int(oracle.field(plan, "grain_seed")[0])
    neg_o, lam_o, cnt_o = oracle.develop_draws(plan, img)
    same = lam_g == lam_o
    n_moved = int(np.count_nonzero(~same))
    # counts at identical means: identical
    assert np.array_equal(cnt_g[same], cnt_o[same]), f"{name}: counts differ at identical Poisson means"
    # every GPU count is the oracle sampler's draw at the GPU's mean
    if n_moved:
        j, y, x = np.nonzero(~same)
        pix = (y * w + x).astype(np.uint32)
        want = oracle.poisson_at(lam_g[j, y, x], pix, STREAMS[j], seed)
        assert np.array_equal(cnt_g[j, y, x], want), f"{name}: counts differ from the sampler at the GPU's means"
        rel = np.abs(lam_g[j, y, x].astype(np.float64) - lam_o[j, y, x]) / np.abs(lam_o[j, y, x])
        assert rel.max() <= 2e-6, f"{name}: Poisson means {rel.max():.3g} apart"
    assert n_moved <= 1e-3 * lam_g.size, f"{name}: {n_moved} of {lam_g.size} Poisson means moved"
    # end to end against the oracle's own chain
    out_o = oracle.print_(plan, neg_o)
    d = np.abs(out.astype(np.int32) - out_o.astype(np.int32)).max(axis=2)
    near = dilate((~same).any(axis=0), 2 + 2)  # the grain blur (radius 2) and the unsharp mask (radius 2)
    far_max = int(d[~near].max()) if (~near).any() else 0
    assert far_max <= 1, f"{name}: {far_max} LSB away from any moved Poisson mean"
    n_over = int(np.count_nonzero(d > 1))
    assert n_over <= 2e-3 * d.size, f"{name}: {n_over} pixels beyond 1 LSB"
    s = dict(name=name, draws=int(lam_g.size), moved_means=n_moved, max_lsb=int(d.max()), px_over_1lsb=n_over)
    print(s, flush=True)
    return s


@pytest.mark.timeout(600)
def test_grain_halation_6000x4000_end_to_end(engine, oracle):
    w, h = 6000, 4000
    print(f"grain+halation {w}x{h}: render", flush=True)
    img = synth.make_frame(w, h, 0)
    d = dict(GRAIN_HALATION, film_format_mm=35.0)
    with engine.open(img, d) as s:
        s.set_overrides(dict(MIX3_OVERRIDES, fir_crossover=16.5, record_grain=True))
        out, _ = s.render("full")
        plan = s.last_plan()
        lam_g, cnt_g = s.grain_record("full")
    assert out.shape == (h, w, 4) and np.all(out[..., 3] == 65535)
    print("grain+halation: oracle", flush=True)
    check_frame(oracle, "grain_halation_6000x4000", plan, img, out, lam_g, cnt_g)


# Frames spanning the batched workload's sizes (log-uniform areas 256x256 .. 4096x2730, both
# orientations): the column passes run at 1365 rows (chunk 64), 2730, 3600 and 4000 rows (chunk 128).
BATCH = [(4096, 2730), (2048, 1365), (1200, 3600), (1024, 683), (640, 640), (256, 256), (400, 1600), (3000, 200),
         (960, 4000)]


@pytest.mark.timeout(600)
def test_batch_matches_oracle_per_frame(engine, oracle, resources):
    import torch
    rng = np.random.default_rng(11)
    sessions, outs, imgs, deltas, overrides = [], [], [], [], []
    try:
        for i, (w, h) in enumerate(BATCH):
            params = dict(GRAIN_HALATION, film_format_mm=35.0, print_exposure=float(rng.uniform(0.5, 2.0)),
                          m_filter_shift=float(rng.uniform(-0.3, 0.3)), y_filter_shift=float(rng.uniform(-0.3, 0.3)))
            mixture = [[sg * float(rng.uniform(0.5, 2.0)), wt] for sg, wt in ((2.0, 0.6), (8.0, 0.3), (32.0, 0.1))]
            ov = {"halation_mixture": mixture, "halation_scatter": False, "fir_crossover": 16.5 if i % 2 else 3.0}
            img = synth.make_frame(w, h, 100 + i)
            s = engine.open(img, params)
            s.set_overrides(ov)
            sessions.append(s)
            imgs.append(img)
            deltas.append(params)
            overrides.append(ov)
            outs.append(torch.zeros((h, w, 4), dtype=torch.int16, device="cuda"))
        print(f"batch of {len(BATCH)} frames, {sum(w * h for w, h in BATCH) / 1e6:.1f} MP", flush=True)
        with spk.Batch(engine, sessions, [o.data_ptr() for o in outs], [w for w, _ in BATCH],
                       {"autotune": False, "record_grain": True}) as b:
            b.run()
            b.sync()
            torch.cuda.synchronize()
            for i, (w, h) in enumerate(BATCH):
                lam_g, cnt_g = b.grain_record(i, w, h)
                # the batch drew each session's first film seed (the seed sequence's first draw);
                # the host plan of the same parameters with that seed is the frame's plan
                out = outs[i].cpu().numpy().view(np.uint16)
                plan = spk.host_plan(resources, deltas[i], w, h, FIRST_FILM_SEED, overrides[i])
                print(f"frame {i} {w}x{h}: oracle", flush=True)
                check_frame(oracle, f"batch_{w}x{h}", plan, imgs[i], out, lam_g, cnt_g)
    finally:
        for s in sessions:
            s.close()


SHAPES = [(64, 1365), (40, 2730), (48, 4000), (6000, 24)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("sigma,crossover", [(8.0, 3.0), (32.0, 16.5)])
def test_recursive_passes_at_bench_heights(engine, oracle, shape, sigma, crossover):
    w, h = shape
    rng = np.random.default_rng(5)
    img = rng.uniform(0.0, 2.0, (h, w, 3)).astype(np.float32)
    sg = (sigma, sigma, sigma)
    gpu = engine.gaussian(img, sg, crossover)
    cpu = oracle.gaussian(img, sg, crossover)
    assert ulps(gpu, cpu).max() <= 2
    gpu_t = engine.gaussian(img, sg, crossover, adjoint=True)
    cpu_t = oracle.gaussian_adjoint(img, sg, crossover)
    assert ulps(gpu_t, cpu_t).max() <= 2
