"""The project's autotuner: times the launch-shape candidates of each knob on the device, per size
class of a workload's pyramid levels, and writes the winners as the table every context loads
(fastwarp_amd/tuning/gfx950.json).

    python -m fastwarp_amd.autotune [--out fastwarp_amd/tuning/gfx950.json] [--repeats 50]

A size class is floor(log2(size)) of a level grid's voxels (CC, smoothing, inversion) or of its
linear-stage samples (MI). The workload is the cohort pyramid (shrink 8/4/2/1) of the 1 mm grid and
the 2 mm grid's levels, on the seeded phantom. Each candidate runs `repeats` times after a warm-up
with every launch timed (HIP events on the context's stream), in three rounds; the knob takes the
candidate with the least total kernel time of the op when it beats the library's default shape by
more than MARGIN, and keeps the default (no entry) otherwise. Choices change results at most by rounding (the CC z runs
change the order of the corrections' sums; block counts change the order of reductions).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

from ._abi import Grid, InvertParams, LinearTransform, MiBinning
from .backend import HipBackend, MiLevel
from .phantom import make_pair

OUT = Path(__file__).resolve().parent / "tuning" / "gfx950.json"
# A candidate must beat the default by this fraction of its time to be kept.
MARGIN = 0.03

CANDIDATES = {
    "cc": [("cc.tile_z", t, "cc.min_blocks", m) for t in (8, 16, 32, 64) for m in (256, 512)],
    "gauss": [("gauss.one_pass", v) for v in (0, 1)],
    "inv": [("inv.max_blocks", v) for v in (512, 1024, 2048, 4096, 8192)],
    "mi": [("mi.threads", t, "mi.max_blocks", b) for t in (256, 512, 1024) for b in (16, 32, 256, 512, 1024, 2048)],
    "mi_deriv": [("mi.deriv_blocks", v) for v in (256, 512, 1024, 2048)],
    "mi_batch": [("mi.batch_blocks", v) for v in (16, 32, 64, 128, 256, 512)],
}


def size_class(n: int) -> int:
    return int(math.floor(math.log2(max(n, 1))))


def level_grid(g: Grid, shrink: int) -> Grid:
    nx, ny, nz = g.shape_xyz
    if shrink == 1:
        return g
    return Grid.make((nx // shrink, ny // shrink, nz // shrink), tuple(s * shrink for s in g.spacing),
                     tuple(o + (shrink - 1) * s / 2 for o, s in zip(g.origin, g.spacing)))


def timed(b: HipBackend, op, families, repeats: int) -> float:
    op()
    b.set_profiling(True)
    for _ in range(repeats):
        op()
    prof = b.take_profile()
    b.set_profiling(False)
    return sum(v[1] for k, v in prof.items() if any(k.startswith(f) for f in families)) / repeats


def apply(b: HipBackend, cand, cls: int, on: bool):
    for q in range(0, len(cand), 2):
        b.set_tuning(f"{cand[q]}@{cls}", cand[q + 1] if on else -1)


def tune(b: HipBackend, name: str, cls: int, op, families, repeats: int, table: dict, log, rounds: int = 3):
    """Times the library's default shape (no entry) and every candidate, `rounds` times in turn,
    each at the least of its rounds; a candidate replaces the default only when it is faster by more than
    `MARGIN`, so that noise between equal shapes leaves no entry."""
    cands = [()] + CANDIDATES[name]
    best_ms = [math.inf] * len(cands)
    for _ in range(rounds):
        for c, cand in enumerate(cands):
            apply(b, cand, cls, True)
            best_ms[c] = min(best_ms[c], timed(b, op, families, repeats))
            apply(b, cand, cls, False)
    for cand, ms in zip(cands, best_ms):
        log(f"  {name}@{cls} {cand or 'default'}: {1e3 * ms:.1f} us")
    c = min(range(len(cands)), key=lambda q: best_ms[q])
    if c and best_ms[c] < (1.0 - MARGIN) * best_ms[0]:
        for q in range(0, len(cands[c]), 2):
            table[f"{cands[c][q]}@{cls}"] = cands[c][q + 1]
        return cands[c]
    return ()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(OUT))
    ap.add_argument("--repeats", type=int, default=50)
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    # start from the library's defaults: the context loads an empty table instead of the shipped one
    os.environ["FASTWARP_TUNING"] = os.devnull
    b = HipBackend(0)
    table: dict = {}
    t0 = time.perf_counter()
    grids = {}
    for shape, spacing, seed in (((182, 218, 182), 1.0, 2), ((96, 112, 96), 2.0, 1)):
        pair = make_pair(shape, spacing, seed=seed)
        for shrink in (8, 4, 2, 1):
            lg = level_grid(pair.grid, shrink)
            grids.setdefault(size_class(lg.n), (pair, lg))
    for cls, (pair, lg) in sorted(grids.items()):
        log(f"size class {cls}: grid {lg.shape_xyz}")
        g = pair.grid
        zero = np.zeros((lg.n, 3))
        f_lvl = b.warp_through_field(b.upload_image(g, pair.fixed[0]), b.upload_field(lg, zero))
        m_lvl = b.warp_through_field(b.upload_image(g, pair.moving[0]), b.upload_field(lg, zero))
        tune(b, "cc", cls, lambda: b.cc_value_derivative_pair(f_lvl, m_lvl, 4), ["k_cc2"], args.repeats, table, log)
        rng = np.random.default_rng(cls)
        i = np.arange(lg.n)
        u = np.stack([np.sin(0.05 * (i % 97)), np.cos(0.03 * (i % 89)), np.sin(0.02 * (i % 83))], 1) * lg.spacing[0]
        du = b.upload_field(lg, u + 0.05 * rng.standard_normal(u.shape))
        tune(b, "gauss", cls, lambda: b.smooth_field(b.copy_field(du), 3.0), ["k_gauss"], args.repeats, table, log)
        no_stop = InvertParams(3, 0.0, 0.0, 1)
        tune(b, "inv", cls, lambda: b.invert(du, b.zero_field(lg), no_stop), ["k_inv"], args.repeats, table, log)
        # the linear stage of this level: every fourth voxel, jittered
        n = lg.n
        idx = np.arange(0, n, 4)
        size = np.array(lg.shape_xyz)
        ijk = np.stack([idx % size[0], (idx // size[0]) % size[1], idx // (size[0] * size[1])], 1).astype(float)
        pts = np.array(list(lg.origin)) + (ijk + rng.normal(0, 1 / 3, ijk.shape)) * np.array(list(lg.spacing))
        binning = MiBinning.from_ranges(32, (0.0, 1.0), (0.0, 1.0))
        fb = rng.integers(2, 30, size=idx.size).astype(np.uint32)
        level = MiLevel(b, m_lvl, pts, fb, binning)
        centre = tuple(np.array(list(lg.origin)) + (size - 1) / 2 * np.array(list(lg.spacing)))
        t = LinearTransform.make(0, [0.01, -0.02, 0.015, 0.5, -0.3, 0.2], centre)
        scls = size_class(idx.size)
        won = tune(b, "mi", scls, lambda: level.value(t), ["k_mi_sampled_hist", "k_mi_sum_slots", "k_mi_finish", "k_mi_sum_finish"],
                   args.repeats, table, log)
        apply(b, won, scls, True)
        tune(b, "mi_deriv", scls, lambda: level.value_and_derivative(t), ["k_mi_sampled_deriv", "k_mi_deriv_finish"],
             args.repeats, table, log)
        # the line search's speculative probes: four transforms in one pass
        probes = [LinearTransform.make(0, [0.01 * q, -0.02, 0.015, 0.5, -0.3 * q, 0.2], centre) for q in range(4)]
        tune(b, "mi_batch", scls, lambda: level.values(probes),
             ["k_mi_sampled_hist_batch", "k_mi_sum_slots", "k_mi_finish", "k_mi_sum_finish"], args.repeats, table, log)
    out = {"_command": "python -m fastwarp_amd.autotune", "_device": b.name(),
           "_seconds": round(time.perf_counter() - t0, 1), **dict(sorted(table.items()))}
    Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    Path(args.out).write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
