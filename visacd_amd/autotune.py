"""Autotuner for the batched frontier step.

Times a step function under every configuration of a search space and
keeps the fastest.  The knobs are the context options of the engine:

* ``lanes_per_pair`` (VACD_OPT_LANES_PER_PAIR): lanes that share one
  depth-first BVH walk of a pair the fan pass left open (1, 2, 4, 8, 16);
* ``front_max`` (VACD_OPT_FRONT_MAX): the longest walk queue walked breadth
  first (8 lanes per pair) instead of depth first (131072; 2^40 = always,
  the library's default);
* ``score_chunk`` (VACD_OPT_SCORE_CHUNK): edges per scoring block, i.e. the
  edge tile a block streams through its 512 planes (0 = sized for about
  8192 blocks; 64 or 256 otherwise).

The traversal stack depth is not a knob: it is sized from the tree depth
(32 or 64 entries), the smallest that cannot overflow.
"""
from __future__ import annotations

import itertools
import time

from visacd_amd import native

SPACE = {
    "lanes_per_pair": (1, 2, 4, 8, 16),
    "front_max": (131072, 1 << 40),
    "score_chunk": (0, 64, 256),
}

_OPTION = {
    "lanes_per_pair": native.VACD_OPT_LANES_PER_PAIR,
    "front_max": native.VACD_OPT_FRONT_MAX,
    "score_chunk": native.VACD_OPT_SCORE_CHUNK,
}

DEFAULT = {"lanes_per_pair": 1, "front_max": 1 << 40, "score_chunk": 0}


def apply(ctx, cfg: dict) -> None:
    """Sets the context options of a configuration."""
    for k, v in cfg.items():
        ctx.set_option(_OPTION[k], int(v))


def tune(ctx, step, sync, space: dict | None = None, repeats: int = 3):
    """Runs step() `repeats` times per configuration (after one untimed run)
    and returns (best configuration, [(configuration, ms per step), ...]).

    sync() must wait for the device work that step() enqueued.  The context
    is left set to the best configuration.
    """
    space = space or SPACE
    keys = list(space)
    table = []
    for values in itertools.product(*(space[k] for k in keys)):
        cfg = dict(zip(keys, values))
        apply(ctx, cfg)
        step()
        sync()
        t0 = time.perf_counter()
        for _ in range(repeats):
            step()
        sync()
        table.append((cfg, (time.perf_counter() - t0) * 1e3 / repeats))
    best = min(table, key=lambda r: r[1])[0]
    apply(ctx, best)
    return best, table
