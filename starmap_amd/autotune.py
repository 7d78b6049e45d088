"""Autotuner for the substep schedules.

For a prepared ``HipSolver`` or ``BatchSolver`` it tries candidate
schedules -- the workgroup tile shape, the LDS budget (workgroups per CU the
schedule is made for) and the number of waves updating the cells (4, or 8
with two per SIMD) -- by re-making the schedules of every handle
(``smp_set_schedule``), timing a few steps with HIP events, and keeping the
fastest.  The state is saved to a device checkpoint before the trials and
restored after each one, so tuning does not advance the solution.  Results
are cached per problem signature (grid, moment count, precision of every
handle) in a JSON file (``SMP_TUNE_CACHE``, default
``~/.cache/starmap_amd/tune.json``).
"""
import hashlib
import json
import os

import numpy as np

TILES = {1: [(255, 1, 1)],
         2: [(63, 4, 1), (127, 2, 1), (31, 8, 1)],
         3: [(15, 4, 4), (31, 4, 2), (15, 8, 2)]}
# "fit": each handle of a batch its own tile, the one with the lowest
# estimated cost for its grid -- tiles x slot size x the cost of a staged
# element for the slot's row length (measured on the 2D P39 1000^2 lattice,
# FP32, two workgroups per CU: rows of 16 elements copy slower than long
# ones) -- so that small grids do not carry a nearly empty row or column of
# tiles
FIT = "fit"
FIT_TILES = {1: [(255, 1, 1)],
             2: [(63, 4, 1), (127, 2, 1), (31, 8, 1), (15, 16, 1)],
             3: [(15, 4, 4), (31, 4, 2), (15, 8, 2)]}
ROW_COST = {16: 1.33, 32: 1.15, 64: 1.0, 128: 0.92, 256: 0.9}
# tiles of up to 512 cells, two neighbouring cells per thread: FP32 state
# only; "fit-pair" picks one of them per handle of a batch
PAIR_TILES = {2: [(63, 8, 1), (127, 4, 1), (255, 2, 1)],
              3: [(31, 4, 4), (15, 8, 4)]}
FIT_PAIR = "fit-pair"
FIT_PAIR_TILES = {2: [(63, 8, 1), (127, 4, 1), (255, 2, 1), (31, 16, 1), (15, 32, 1)],
                  3: [(31, 4, 4), (15, 8, 4)]}
BLOCKS = (1, 2)
CELL_WAVES = (4, 8)
_SLOT = 4095            # checkpoint slot used while tuning


def _solvers(target):
    return target.solvers if hasattr(target, "solvers") else [target]


def signature(target):
    items = []
    for s in _solvers(target):
        items.append([int(v) for v in s.n] + [int(s.n_sys), s.precision, int(len(s.gkey))])
    return hashlib.sha1(json.dumps([type(target).__name__, items]).encode()).hexdigest()


def _cache_path():
    return os.environ.get("SMP_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "starmap_amd",
                                       "tune.json"))


def _load_cache():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(key, entry):
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        c = _load_cache()
        c[key] = entry
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(c, f, indent=1)
        os.replace(tmp, path)
    except OSError:
        pass


def _dim(s):
    return max(int(np.sum(np.asarray(s.n) > 1)), 1)


def candidates(target):
    dim = max(_dim(s) for s in _solvers(target))
    tiles = list(TILES[dim])
    if hasattr(target, "solvers"):
        tiles.append(FIT)
    out = [(t, b, c) for t in tiles for b in BLOCKS for c in CELL_WAVES]
    if dim > 1 and all(s.precision == "fp32" for s in _solvers(target)):
        pt = list(PAIR_TILES[dim]) + ([FIT_PAIR] if hasattr(target, "solvers") else [])
        out += [(t, b, c) for t in pt for b in BLOCKS for c in CELL_WAVES]
    return out


def staged_elements(s, tile):
    """Elements one substep stages for handle s with this tile: tiles times
    the slot size (tile + one-cell halo per axis)."""
    ext = [int(v) for v in np.asarray(s.n) if v > 1] or [1]
    tiles = 1
    slot = 1
    for e, n in enumerate(ext):
        tiles *= -(-n // tile[e])
        slot *= tile[e] + 1
    return tiles * slot


def fit_tiles(s, pair=False):
    """FIT_TILES (or FIT_PAIR_TILES) of the handle's dimension, lowest
    estimated cost first."""
    table = FIT_PAIR_TILES if pair else FIT_TILES
    return sorted(table[_dim(s)], key=lambda t: staged_elements(s, t) * ROW_COST[t[0] + 1])


def apply(target, tile, blocks_per_cu, cell_waves):
    for s in _solvers(target):
        if tile not in (FIT, FIT_PAIR):
            s.set_schedule(tile, blocks_per_cu, cell_waves)
            continue
        err = None
        for t in fit_tiles(s, pair=tile == FIT_PAIR):
            try:
                s.set_schedule(t, blocks_per_cu, cell_waves)
                err = None
                break
            except Exception as e:       # a tile the coupling does not fit
                err = e
        if err is not None:
            raise err


def _time(target, steps):
    """Kernel milliseconds of `steps` steps (HIP events)."""
    if hasattr(target, "solvers"):
        target.step(np.full(len(target), steps), use_graph=False)
        return target.timing()[1]
    target.step(target.dt_prepared, steps)
    return target.timing()[1]


def tune(target, steps=2, use_cache=True, verbose=False):
    """Choose and apply the fastest schedule for a prepared target; returns
    ((tile, blocks_per_cu, cell_waves), {candidate: ms})."""
    key = signature(target)
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            tile = hit["tile"] if hit["tile"] in (FIT, FIT_PAIR) else tuple(hit["tile"])
            best = (tile, hit["blocks_per_cu"], hit["cell_waves"])
            apply(target, *best)
            return best, {}
    sol = _solvers(target)
    for s in sol:
        s.checkpoint_save(_SLOT)
    times = {}
    try:
        for cand in candidates(target):
            try:
                apply(target, *cand)
            except Exception as e:        # a schedule the coupling does not fit
                if verbose:
                    print("autotune: skip", cand, e)
                continue
            _time(target, 1)             # first use: coefficient upload, warm-up
            for s in sol:
                s.checkpoint_load(_SLOT)
            ms = _time(target, steps)
            for s in sol:
                s.checkpoint_load(_SLOT)
            times[cand] = ms
            if verbose:
                print("autotune:", cand, "%.3f ms" % ms)
    finally:
        for s in sol:
            s.checkpoint_load(_SLOT)
            s.checkpoint_free(_SLOT)
    if not times:
        raise RuntimeError("autotune: no candidate schedule fits")
    best, _ = min(times.items(), key=lambda kv: kv[1])
    apply(target, *best)
    _store_cache(key, {"tile": best[0] if best[0] in (FIT, FIT_PAIR) else list(best[0]),
                       "blocks_per_cu": best[1], "cell_waves": best[2],
                       "ms": {"%s/%d/%d" % c: v for c, v in times.items()}})
    return best, times
