"""Launch-shape autotuner (measured, not modelled).

``tune_trace`` picks the persistent grid size (workgroups per CU) of
``mx_trace`` for one compiled instrument and photon count; ``tune_batched``
picks the tile plan of a :class:`marxs_amd.batch.BatchedTrace`: photons per
tile (one size for all variants, or sizes scaled by each variant's estimated
tests per photon), queue order and grid size.  Every candidate is timed with
HIP events on the launch stream (median of ``reps`` launches after one
warm-up); the fastest is applied and returned together with the table of
timings.  Results are cached per process by a key of the problem's shape.
"""
import torch

__all__ = ['tune_trace', 'tune_batched', 'time_launch', 'CACHE']

CACHE = {}


def time_launch(fn, reps=3, device=None):
    """Median milliseconds of ``fn()`` (which launches on the current
    stream), measured with events on that stream."""
    stream = torch.cuda.current_stream(device)
    fn()
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fn()
        b.record(stream)
        b.synchronize()
        times.append(a.elapsed_time(b))
    times.sort()
    return times[len(times) // 2]


def tune_trace(ci, dp, uniforms, precision=32, candidates=(1, 2, 3, 4), reps=5, src=None):
    """Workgroups per CU for ``ci.launch`` on these photons; sets
    ``ci.wg_per_cu[precision]`` and returns (best, {candidate: ms}).  The
    candidates are the grids that came within a few per cent of each other on
    the 600-facet workloads (which of 1 and 2 per CU is faster differed from
    one MI355X to another by up to 3.5 %)."""
    key = ('trace', id(ci.tables), dp.n, int(precision))
    if key in CACHE:
        ci.wg_per_cu[int(precision)] = CACHE[key][0]
        return CACHE[key]
    status = torch.zeros(8, dtype=torch.int64, device=ci.device)
    res = {}
    for c in candidates:
        res[c] = time_launch(lambda: ci.launch(dp, uniforms, precision, status=status,
                                               src=src, wg_per_cu=c), reps, ci.device)
    best = min(res, key=res.get)
    ci.wg_per_cu[int(precision)] = best
    CACHE[key] = (best, res)
    return best, res


def _batched_candidates(bt):
    from .batch import BLOCK, BLOCK_EXT
    bs = BLOCK_EXT if bt.ext_stages else BLOCK[bt.precision]
    out = []
    for mult in (1, 2, 4, 8):
        for scaled in (False, True):
            out.append({'tile_photons': bs * mult, 'scale_by_cost': scaled, 'order': 'cost',
                        'wg_per_cu': 1})
    out.append({'tile_photons': bs * 2, 'scale_by_cost': False, 'order': 'given', 'wg_per_cu': 1})
    out.append({'tile_photons': bs * 2, 'scale_by_cost': False, 'order': 'cost', 'wg_per_cu': 2})
    return out


def tune_batched(bt, candidates=None, reps=3):
    """Tile plan for a BatchedTrace; applies the fastest and returns
    (plan arguments, [(arguments, ms), ...])."""
    key = ('batched', tuple(len(ci.tables.elements) for ci in bt.instruments),
           tuple(src.n for src in bt.inputs), bt.precision)
    if candidates is None and key in CACHE:
        best = CACHE[key][0]
        bt.set_plan(**best)
        return CACHE[key]
    cands = _batched_candidates(bt) if candidates is None else list(candidates)
    table = []
    for c in cands:
        bt.set_plan(**c)
        table.append((c, time_launch(bt.launch, reps, bt.device)))
    best = min(table, key=lambda r: r[1])[0]
    bt.set_plan(**best)
    if candidates is None:
        CACHE[key] = (best, table)
    return best, table
