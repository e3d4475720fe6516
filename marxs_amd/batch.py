"""Many instrument variants traced in one persistent launch.

The reference evaluates design variants one after the other
(``run_tolerances``, ``marxs/design/tolerancing.py:224-283``: modify the
instrument, trace ``photons_in.copy()``, analyse).  :class:`BatchedTrace`
traces all variants of a study in a single ``mx_trace_batched`` launch: each
variant keeps its own compiled tables (ragged facet counts) and its own
photons, the work is split into tiles of photons, and a device-side queue
hands the tiles to persistent workgroups, which restage their LDS tables only
when the variant changes.  Every variant's result is the one
``CompiledInstrument.launch`` gives for it alone.

The launch (queue reset, status reset and the kernel) can be captured once in
a hipGraph (:meth:`BatchedTrace.capture`) and replayed; inputs are read from
separate input columns, so replays are repeatable.

Tile plans (photons per tile for each variant, queue order, workgroups per
CU) come from :func:`marxs_amd.autotune.tune_batched`.
"""
import copy
import ctypes

import numpy as np
import torch

from . import _lib
from .compiler import _build_lds
from .tracer import (CompiledInstrument, DevicePhotons, DeviceTables, PREC, raise_for_status,
                     TraceStatus, ST_PROB_RANGE, ST_ORDER_EMPTY, ST_SLOT_REUSE,
                     ST_GRID_FALLBACK, ST_MULTI_HIT, ST_TRANS_RANGE, ST_TESTS)

__all__ = ['BatchedTrace', 'variant_cost', 'outputs_like']

BLOCK = {32: 1024, 64: 768}
BLOCK_EXT = 768      # work-items per workgroup of the kernels with stack or tube stages


def outputs_like(src):
    """Output photon columns for a trace that reads ``src`` (energy shared)."""
    dp = DevicePhotons.__new__(DevicePhotons)
    dp.device, dp.n = src.device, src.n
    f = dict(dtype=torch.float64, device=src.device)
    dp.pos = torch.empty((3, src.n), **f)
    dp.dir = torch.empty((3, src.n), **f)
    dp.pol = torch.empty((3, src.n), **f)
    dp.prob = torch.empty(src.n, **f)
    dp.energy = src.energy
    dp.has_pos = src.has_pos
    dp.fcols, dp.icols = {}, {}
    return dp


def variant_cost(ci):
    """Estimated element tests per photon of a compiled instrument: mean
    candidate-list length of the culling cells of each gridded stage, the
    element count of the others."""
    t = ci.tables
    cost = 0.0
    for st in t.stages:
        n = len(st.elements)
        g = st.grid
        if g is None:
            cost += n
        else:
            lens = np.diff(np.asarray(g['cell_start'], dtype=np.int64))
            nz = lens[lens > 0]
            cost += float(nz.mean()) if len(nz) else 1.0
    return cost


def _tables_with_budget(ci, budget):
    """Device tables of ``ci`` whose LDS image holds at most ``budget`` bytes."""
    cache = ci.__dict__.setdefault('_budget_tables', {})
    if budget not in cache:
        t = copy.copy(ci.tables)
        t.stage_i = ci.tables.stage_i.copy()
        t.sel_i = ci.tables.sel_i.copy()
        _build_lds(t, budget, budget)
        cache[budget] = DeviceTables(t, ci.device)
    return cache[budget]


class BatchedTrace:
    """All variants of a study in one launch.

    Parameters
    ----------
    instruments : list of CompiledInstrument (one device)
    inputs : list of DevicePhotons, the photons entering each variant (read
        only; results go to ``self.outputs``)
    uniforms : list of (slots, N) float64 tensors (or None for variants that
        draw nothing)
    precision : 64 (parity path) or 32 (FP32 local frame)
    """

    def __init__(self, instruments, inputs, uniforms=None, precision=32):
        if len(instruments) != len(inputs):
            raise ValueError('one input photon set per instrument')
        if len(instruments) == 0:
            raise ValueError('no variants')
        self.instruments = list(instruments)
        self.inputs = list(inputs)
        self.uniforms = list(uniforms) if uniforms is not None else [None] * len(inputs)
        self.precision = PREC[precision]
        self.device = self.instruments[0].device
        for ci in self.instruments:
            if ci.device != self.device:
                raise ValueError('all variants must live on one device')
            if ci.tables.required_icols:
                # e.g. a QualityFactor outside a stack reads the photons'
                # order column, which the batched outputs do not carry in
                raise ValueError('batched traces start without photon columns; a variant reads '
                                 + ', '.join(ci.tables.required_icols))
        self.n_variants = len(self.instruments)
        self.outputs = [outputs_like(src) for src in self.inputs]
        self._photon_structs = [ci.photons_struct(dp, u, src) for ci, dp, u, src in
                                zip(self.instruments, self.outputs, self.uniforms, self.inputs)]
        self.costs = np.array([variant_cost(ci) for ci in self.instruments])
        self.ext_stages = any(ci.dev_tables.struct.ext_stages for ci in self.instruments)
        self.queue = torch.zeros(2, dtype=torch.int32, device=self.device)
        self.graph = None
        self.plan = None
        self.set_plan()

    # ------------------------------------------------------------ plan
    def set_plan(self, tile_photons=None, scale_by_cost=False, order='cost', wg_per_cu=1,
                 lds_budget=None, grid=True):
        """Choose the tiling.

        tile_photons: photons per tile (int, or one per variant); default one
        workgroup's worth.  scale_by_cost: shrink tiles of expensive variants
        so that tiles take similar times.  order: 'cost' (most expensive
        variants first), 'given', or a permutation.  lds_budget: bytes of
        tables staged in LDS per workgroup (None = compile default).
        """
        bs = BLOCK_EXT if self.ext_stages else BLOCK[self.precision]
        nv = self.n_variants
        if tile_photons is None:
            tile_photons = bs
        tp = np.broadcast_to(np.asarray(tile_photons, dtype=np.int64), (nv,)).copy()
        if scale_by_cost:
            ref = float(np.median(self.costs))
            tp = np.maximum(bs, np.round(tp * ref / self.costs / bs) * bs).astype(np.int64)
        if np.any(tp < 1) or np.any(tp > 2 ** 31 - 1):
            raise ValueError('photons per tile out of range')
        if isinstance(order, str):
            if order == 'cost':
                perm = np.argsort(-self.costs * np.array([src.n for src in self.inputs]),
                                  kind='stable')
            elif order == 'given':
                perm = np.arange(nv)
            else:
                raise ValueError(f'unknown order {order!r}')
        else:
            perm = np.asarray(order, dtype=np.int64)
            if sorted(perm.tolist()) != list(range(nv)):
                raise ValueError('order must be a permutation of the variants')
        if lds_budget is None:
            dts = [ci.dev_tables for ci in self.instruments]
        else:
            dts = [_tables_with_budget(ci, int(lds_budget)) for ci in self.instruments]
        n = np.array([src.n for src in self.inputs], dtype=np.int64)
        tiles = (n + tp - 1) // tp
        tile_first = np.zeros(nv + 1, dtype=np.int64)
        tile_first[1:] = np.cumsum(tiles[perm])
        # host struct arrays in queue order, and their device copies
        TA = _lib.MxTables * nv
        PA = _lib.MxPhotons * nv
        self._tables_host = TA(*[dts[v].struct for v in perm])
        self._photons_host = PA(*[self._photon_structs[v] for v in perm])
        self._tables_dev = torch.frombuffer(bytearray(bytes(self._tables_host)),
                                            dtype=torch.uint8).to(self.device)
        self._photons_dev = torch.frombuffer(bytearray(bytes(self._photons_host)),
                                             dtype=torch.uint8).to(self.device)
        self._tile_first = torch.as_tensor(tile_first, device=self.device)
        self._tile_photons = torch.as_tensor(tp[perm].astype(np.int32), device=self.device)
        self._dts = dts
        self._perm = perm
        self._status_q = torch.zeros((nv, _lib.MX_NSTATUS), dtype=torch.int64, device=self.device)
        b = _lib.MxBatch()
        b.n_variants = nv
        b.grid = 0
        b.n_tiles = int(tile_first[-1])
        b.tables = ctypes.cast(self._tables_host, ctypes.c_void_p)
        b.photons = ctypes.cast(self._photons_host, ctypes.c_void_p)
        b.tables_dev = self._tables_dev.data_ptr()
        b.photons_dev = self._photons_dev.data_ptr()
        b.tile_first = self._tile_first.data_ptr()
        b.tile_photons = self._tile_photons.data_ptr()
        b.queue = self.queue.data_ptr()
        b.status = self._status_q.data_ptr()
        self._batch = b
        self._flags = (0 if grid else 1) | ((int(wg_per_cu) & 0xff) << 8)
        self.plan = {'tile_photons': tp.tolist(), 'scale_by_cost': bool(scale_by_cost),
                     'order': perm.tolist(), 'wg_per_cu': int(wg_per_cu),
                     'lds_budget': lds_budget, 'n_tiles': int(tile_first[-1])}
        self.graph = None
        return self.plan

    # ---------------------------------------------------------- launch
    def launch(self, stream=None):
        """Reset the status counters and launch the batch (asynchronous)."""
        if stream is None:
            stream = torch.cuda.current_stream(self.device).cuda_stream
        self._status_q.zero_()
        _lib.check(_lib.lib.mx_trace_batched(ctypes.byref(self._batch), self.precision,
                                             self._flags, ctypes.c_void_p(stream)))

    def capture(self):
        """Capture the launch in a hipGraph (torch.cuda.CUDAGraph)."""
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            self.launch(s.cuda_stream)      # warm-up outside the graph
        torch.cuda.current_stream(self.device).wait_stream(s)
        torch.cuda.synchronize(self.device)
        with torch.cuda.graph(g, stream=s):
            self.launch(s.cuda_stream)
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()

    def run(self, use_graph=False):
        """Launch (or replay the graph), wait and check every variant's
        status; returns one TraceStatus per variant (input order)."""
        if use_graph:
            self.replay()
        else:
            self.launch()
        raw = self.raw_status()
        for r in raw:
            raise_for_status(r)
        for dp in self.outputs:
            dp.has_pos = True
        return [TraceStatus(prob_range=int(r[ST_PROB_RANGE]), order_empty=int(r[ST_ORDER_EMPTY]),
                            slot_reuse=int(r[ST_SLOT_REUSE]), grid_fallback=int(r[ST_GRID_FALLBACK]),
                            multi_hit=int(r[ST_MULTI_HIT]), trans_range=int(r[ST_TRANS_RANGE]),
                            tests=int(r[ST_TESTS]))
                for r in raw]

    def raw_status(self):
        """(n_variants, MX_NSTATUS) counters of the last launch, input order."""
        q = self._status_q.cpu().numpy()
        out = np.empty_like(q)
        out[self._perm] = q
        return out
