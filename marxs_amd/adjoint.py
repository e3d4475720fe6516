"""Gradients of a detector-plane line centroid and width with respect to the
6-DOF poses and periods of grating facets (pathwise adjoint).

The reference has no derivatives: it estimates alignment tolerances by
re-tracing perturbed instruments (``run_tolerances``,
``marxs/design/tolerancing.py:224-283``) with facet perturbations
``compose(t, euler2mat(rx, ry, rz, 'sxyz'), 1)`` (``uncertainties.py:36-38``)
inserted into the ``Parallel`` pose chain (``simulator.py:423-437``).  Here
the same perturbations are parameters ``theta[f] = (dx, dy, dz, rx, ry, rz)``
plus the period ``D[f]``; with every photon's facet, order and CCD frozen
(pathwise derivative) the loss

    L = mu + sigma,   mu = sum w y / W,   sigma^2 = sum w (y - mu)^2 / W

of the global detector coordinate ``y`` of the selected photons is smooth in
them.  ``facet_geometry`` builds the facet frames from the parameters with
torch (differentiable); :class:`FacetLoss` evaluates ``L`` with the HIP
replay kernel and its backward runs the HIP adjoint kernel
(``mx_adjoint``), which recomputes every intersection instead of storing it.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from .mathutils import decompose44
from .compiler import compile_tree, STAGE_GRATING, STAGE_DETECTOR

__all__ = ['facet_geometry', 'FacetLoss', 'AdjointProblem', 'ChunkedAdjoint', 'entrance_chunk',
           'euler_sxyz_torch', 'central_differences']


def _rot(axis, a):
    c, s = torch.cos(a), torch.sin(a)
    o, z = torch.ones_like(a), torch.zeros_like(a)
    if axis == 0:
        rows = [[o, z, z], [z, c, -s], [z, s, c]]
    elif axis == 1:
        rows = [[c, z, s], [z, o, z], [-s, z, c]]
    else:
        rows = [[c, -s, z], [s, c, z], [z, z, o]]
    return torch.stack([torch.stack(r, dim=-1) for r in rows], dim=-2)


def euler_sxyz_torch(r):
    """(..., 3) angles -> (..., 3, 3): fixed-axis rotation x, then y, then z."""
    return _rot(2, r[..., 2]) @ _rot(1, r[..., 1]) @ _rot(0, r[..., 0])


def facet_geometry(base, theta, period, groove, sign_mode, transmission):
    """Facet geometry records (F, 16) from pose parameters.

    base: dict of float64 tensors from :meth:`AdjointProblem.pose_base`:
    ``GU`` (4, 4) = container pose . uncertainty, ``t0`` (F, 3) = element
    centre + common offset, ``Rpost`` (F, 3, 3) = element rotation . common
    rotation, ``zoom`` (F, 3).  theta: (F, 6) = translation, sxyz angles of
    the per-element uncertainty.  period: (F,).
    """
    GU = base['GU']
    t = theta[:, :3] + base['t0']
    rot = euler_sxyz_torch(theta[:, 3:6]) @ base['Rpost']
    lin = GU[:3, :3] @ (rot * base['zoom'][:, None, :])
    center = (GU[:3, :3] @ t[..., None])[..., 0] + GU[:3, 3]
    ex = lin[:, :, 0] / torch.linalg.norm(lin[:, :, 0], dim=1, keepdim=True)
    ey = lin[:, :, 1] / torch.linalg.norm(lin[:, :, 1], dim=1, keepdim=True)
    ez = lin[:, :, 2] / torch.linalg.norm(lin[:, :, 2], dim=1, keepdim=True)
    extra = torch.stack([groove, sign_mode, transmission], dim=1)
    return torch.cat([center, ex, ey, ez, period[:, None], extra], dim=1)


def _launch(args, stream):
    _lib.check(_lib.lib.mx_adjoint(ctypes.byref(args), ctypes.c_void_p(stream)))


class AdjointProblem:
    """Frozen photon paths through one facet array and one detector array.

    Built from a traced instrument: ``entrance`` holds the photons as they
    reach the grating stage (pos, dir: (3, N) float64 tensors), ``facet``,
    ``order``, ``ccd`` the frozen choices (int32, -1 = not used), ``weight``
    the photon weights (0 excludes a photon), ``energy`` (keV).
    """

    def __init__(self, gratings, detectors, pos, dir, energy, weight, facet, order, ccd,
                 coord=1, device=None):
        self.device = torch.device(device) if device is not None else pos.device
        self.gratings = gratings
        self.facets = list(gratings.elements)
        from .optics import FlatGrating, FlatDetector
        from .compiler import CompileError
        if not all(isinstance(e, FlatGrating) and not callable(e._d) for e in self.facets):
            raise CompileError('the adjoint replays flat grating facets with a fixed period '
                               '(FlatGrating, CATGrating, NonParallelCATGrating); stacks and '
                               'callable periods are not differentiated')
        dets = list(detectors.elements) if hasattr(detectors, 'elements') else list(detectors)
        if not all(isinstance(d, FlatDetector) for d in dets):
            raise CompileError('the adjoint replays flat detectors (FlatDetector) only')
        self.coord = coord
        f64 = dict(dtype=torch.float64, device=self.device)
        self.pos = pos.to(**f64).contiguous()
        self.dir = dir.to(**f64).contiguous()
        self.energy = energy.to(**f64).contiguous()
        self.weight = weight.to(**f64).contiguous()
        self.facet = facet.to(dtype=torch.int32, device=self.device).contiguous()
        self.order = order.to(dtype=torch.int32, device=self.device).contiguous()
        self.ccd = ccd.to(dtype=torch.int32, device=self.device).contiguous()
        self.n = self.pos.shape[1]
        ct = compile_tree(list(detectors.elements) if hasattr(detectors, 'elements') else detectors)
        self.ccd_t = torch.as_tensor(ct.elem_t, **f64).contiguous()
        self.n_ccd = self.ccd_t.shape[0]
        F = len(self.facets)
        self.groove = torch.tensor([e.geometry['groove_angle'] for e in self.facets], **f64)
        self.sign_mode = torch.tensor([1. if e.sign_convention == 'cat' else 0. for e in self.facets], **f64)
        self.transmission = torch.tensor([1. if e.transmission else 0. for e in self.facets], **f64)
        self.period0 = torch.tensor([float(e._d) for e in self.facets], **f64)
        self.base = self.pose_base()
        self.theta0 = self.current_theta()
        self.F = F

    # ------------------------------------------------------------ poses
    def pose_base(self):
        """Fixed parts of the element pose chain (simulator.py:423-437)."""
        g = self.gratings
        f64 = dict(dtype=torch.float64, device=self.device)
        GU = torch.as_tensor(g.pos4d @ g.uncertainty, **f64)
        t0, rpost, zoom = [], [], []
        for i, e in enumerate(self.facets):
            from .base import parse_position_keywords
            args = g._specific_args(i)
            ep = parse_position_keywords(args)
            te, re, ze, _ = decompose44(ep)
            tc, rc, zc, _ = decompose44(g.elem_pos[i])
            t0.append(tc + te)
            rpost.append(rc @ re)
            zoom.append(zc * ze)
        return {'GU': GU, 't0': torch.as_tensor(np.array(t0), **f64),
                'Rpost': torch.as_tensor(np.array(rpost), **f64),
                'zoom': torch.as_tensor(np.array(zoom), **f64)}

    def current_theta(self):
        """6-DOF parameters of the current elem_uncertainty (translation and
        sxyz angles recovered from each 4x4 matrix)."""
        th = []
        for u in self.gratings.elem_uncertainty:
            t, r, z, s = decompose44(u)
            # sxyz: R = Rz(c) Ry(b) Rx(a)
            b = -math.asin(max(-1., min(1., r[2, 0])))
            a = math.atan2(r[2, 1], r[2, 2])
            c = math.atan2(r[1, 0], r[0, 0])
            th.append([t[0], t[1], t[2], a, b, c])
        return torch.tensor(th, dtype=torch.float64, device=self.device)

    def geometry(self, theta, period):
        return facet_geometry(self.base, theta, period, self.groove, self.sign_mode,
                              self.transmission)

    # ----------------------------------------------------------- kernels
    def _args(self, geom, precision, seed=None, y_out=None, grad=None, moments=None, src=None):
        """Kernel arguments over all photons, or over the compacted photon
        columns ``src`` (a dict from :meth:`_pose_buffers`)."""
        c = src if src is not None else {
            'n': self.n, 'pos': self.pos, 'dir': self.dir, 'energy': self.energy,
            'weight': self.weight, 'facet': self.facet, 'order': self.order, 'ccd': self.ccd}
        a = _lib.MxAdjointArgs()
        a.n = c['n']
        for k in range(3):
            a.pos[k] = c['pos'][k].data_ptr()
            a.dir[k] = c['dir'][k].data_ptr()
        a.energy = c['energy'].data_ptr()
        a.weight = c['weight'].data_ptr()
        a.facet = c['facet'].data_ptr()
        a.order = c['order'].data_ptr()
        a.ccd = c['ccd'].data_ptr()
        a.geom = geom.data_ptr()
        a.ccd_t = self.ccd_t.data_ptr()
        a.seed = seed.data_ptr() if seed is not None else None
        a.y_out = y_out.data_ptr() if y_out is not None else None
        a.grad = grad.data_ptr() if grad is not None else None
        a.moments = moments.data_ptr() if moments is not None else None
        a.n_facet = geom.shape[0]
        a.n_ccd = self.n_ccd
        a.coord = self.coord
        a.precision = {64: 64, 32: 32}[precision]
        a.lds_accumulate = 1 if geom.shape[0] * 13 * 8 <= 64 * 1024 else 0
        return a

    def coordinates(self, geom, precision=64):
        """Per-photon detector coordinate (NaN for unused photons)."""
        geom = geom.detach().contiguous()
        y = torch.empty(self.n, dtype=torch.float64, device=self.device)
        _launch(self._args(geom, precision, y_out=y), torch.cuda.current_stream(self.device).cuda_stream)
        return y

    def backward_kernel(self, geom, seed, precision=64):
        geom = geom.detach().contiguous()
        grad = torch.zeros((geom.shape[0], 13), dtype=torch.float64, device=self.device)
        _launch(self._args(geom, precision, seed=seed.contiguous(), grad=grad),
                torch.cuda.current_stream(self.device).cuda_stream)
        return grad

    def loss(self, theta, period, precision=64):
        """L(theta, period) = centroid + RMS width (differentiable)."""
        geom = self.geometry(theta, period)
        return FacetLoss.apply(geom, self, precision)

    # -------------------------------------------------- fused fast path
    def _pose_buffers(self):
        """Device copies of the fixed pose-chain parts and the work buffers of
        :meth:`loss_and_grad` (allocated once, so that the call can be
        captured in a hipGraph)."""
        if getattr(self, '_pb', None) is not None:
            return self._pb
        f64 = dict(dtype=torch.float64, device=self.device)
        b = self.base
        F = self.F
        pb = {'gu': b['GU'].contiguous().reshape(16).clone(),
              't0': b['t0'].contiguous().clone(),
              'rpost': b['Rpost'].contiguous().reshape(F, 9).clone(),
              'zoom': b['zoom'].contiguous().clone(),
              'extra': torch.stack([self.groove, self.sign_mode, self.transmission], 1).contiguous(),
              'theta': torch.zeros((F, 6), **f64), 'period': torch.zeros(F, **f64),
              'geom': torch.zeros((F, 16), **f64), 'grad7': torch.zeros((F, 7), **f64),
              'loss': torch.zeros(1, **f64)}
        # the single pass accumulates per facet the sums A, B (26 words) and the
        # three moments: one buffer, cleared by one fill
        pb['accum'] = torch.zeros(F * 26 + 3, **f64)
        pb['sums'] = pb['accum'][:F * 26].view(F, 26)
        pb['moments'] = pb['accum'][F * 26:]
        # shift of the moments: the weighted mean coordinate at the base pose
        y = self.coordinates(self.geometry(self.theta0, self.period0))
        ok = ~torch.isnan(y)
        w = torch.where(ok, self.weight, torch.zeros_like(self.weight))
        self.y_shift = float((w * torch.where(ok, y, torch.zeros_like(y))).sum() / w.sum())
        # stream compaction: the fused kernels visit only the photons that
        # carry weight on a facet and a detector
        idx = torch.nonzero((self.weight != 0) & (self.facet >= 0) & (self.ccd >= 0)).flatten()
        pb['src'] = {'n': int(idx.numel()), 'idx': idx,
                     'pos': self.pos[:, idx].contiguous(), 'dir': self.dir[:, idx].contiguous(),
                     'energy': self.energy[idx].contiguous(), 'weight': self.weight[idx].contiguous(),
                     'facet': self.facet[idx].contiguous(), 'order': self.order[idx].contiguous(),
                     'ccd': self.ccd[idx].contiguous()}
        self._pb = pb
        return pb

    def restrict_to_traced(self, y_traced, rtol=1e-10):
        """Zero the weight of photons whose traced detector coordinate the
        one-facet replay does not reproduce (photons that met more than one
        facet or detector in the trace).  Returns the number excluded.

        The work buffers of :meth:`loss_and_grad` are rebuilt afterwards; a
        graph captured before keeps its own buffers alive (they travel with
        the graph) and replays the old photon set: capture again."""
        self._pb = None
        y = self.coordinates(self.geometry(self.theta0, self.period0))
        used = self.weight != 0
        bad = used & ~(torch.abs(y - y_traced) <= rtol * torch.abs(y_traced).clamp(min=1.0))
        self.weight = torch.where(bad, torch.zeros_like(self.weight), self.weight).contiguous()
        return int(bad.sum())

    def _pose_args(self, mode):
        pb = self._pb
        a = _lib.MxPoseArgs()
        a.n_facet = self.F
        a.mode = mode
        for k in ('theta', 'period', 'gu', 't0', 'rpost', 'zoom', 'extra', 'geom', 'grad7',
                  'moments', 'loss'):
            setattr(a, k, pb[k].data_ptr())
        a.grad13 = pb['accum'].data_ptr()
        a.y_shift = self.y_shift
        return a

    def launch_loss_and_grad(self, precision=32, stream=None):
        """Launch the fused evaluation on the parameters held in
        ``buffers['theta']``/``['period']`` (asynchronous): one fill of the
        accumulators, the pose chain, one replay of every photon that
        accumulates the moments and the per-facet sums A = sum w dy/dg,
        B = sum w (y - s) dy/dg (``MX_AJ_SUMS``), then the chain rule back to
        the poses with dL/dg = A / W + (B - a A) / (W sigma) (L = mu + sigma,
        dL/dy_i = w_i / W (1 + (y_i - mu) / sigma)).  Results:
        ``buffers['loss']`` and ``buffers['grad7']`` (d/d theta, d/d period)."""
        pb = self._pose_buffers()
        if stream is None:
            stream = torch.cuda.current_stream(self.device).cuda_stream
        pb['accum'].zero_()
        _lib.check(_lib.lib.mx_facet_pose(ctypes.byref(self._pose_args(0)), ctypes.c_void_p(stream)))
        one = self._args(pb['geom'], precision, grad=pb['accum'], moments=pb['moments'], src=pb['src'])
        one.flags = _lib.MX_AJ_SUMS
        one.lds_accumulate = int(self.F <= 315)
        one.y_shift = self.y_shift
        _launch(one, stream)
        _lib.check(_lib.lib.mx_facet_pose(ctypes.byref(self._pose_args(2)), ctypes.c_void_p(stream)))

    @property
    def buffers(self):
        return self._pose_buffers()

    def loss_and_grad(self, theta, period, precision=32):
        """L = centroid + RMS width and its gradient with respect to the 6-DOF
        poses and periods, without autograd: returns (L, dL/dtheta (F, 6),
        dL/dperiod (F,)), all on the device."""
        pb = self._pose_buffers()
        pb['theta'].copy_(theta)
        pb['period'].copy_(period)
        self.launch_loss_and_grad(precision)
        return pb['loss'][0].clone(), pb['grad7'][:, :6].clone(), pb['grad7'][:, 6].clone()

    def capture(self, precision=32):
        """hipGraph of :meth:`launch_loss_and_grad` (replay after writing new
        parameters into ``buffers['theta']``/``['period']``)."""
        self._pose_buffers()
        torch.cuda.synchronize(self.device)
        s = torch.cuda.Stream(self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            self.launch_loss_and_grad(precision, s.cuda_stream)
        torch.cuda.current_stream(self.device).wait_stream(s)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            self.launch_loss_and_grad(precision, s.cuda_stream)
        # the graph holds raw pointers into these buffers: keep them alive for
        # as long as the graph, whatever later happens to self._pb
        g._marxs_buffers = self._pb
        return g


def _moments(y, w):
    ok = ~torch.isnan(y)
    wv = torch.where(ok, w, torch.zeros_like(w))
    yv = torch.where(ok, y, torch.zeros_like(y))
    W = wv.sum()
    mu = (wv * yv).sum() / W
    var = (wv * (yv - mu) ** 2).sum() / W
    return W, mu, torch.sqrt(var), ok


class FacetLoss(torch.autograd.Function):
    """L = mu + sigma of the detector coordinate; backward = HIP adjoint."""

    @staticmethod
    def forward(ctx, geom, prob, precision):
        y = prob.coordinates(geom, precision)
        W, mu, sigma, ok = _moments(y, prob.weight)
        ctx.prob = prob
        ctx.precision = precision
        ctx.save_for_backward(geom.detach(), y, mu, sigma, W)
        return mu + sigma

    @staticmethod
    def backward(ctx, gL):
        geom, y, mu, sigma, W = ctx.saved_tensors
        prob = ctx.prob
        ok = ~torch.isnan(y)
        seed = torch.where(ok, prob.weight / W * (1 + (y - mu) / sigma), torch.zeros_like(y))
        g = prob.backward_kernel(geom, seed * gL, ctx.precision)
        full = torch.zeros_like(geom)
        full[:, :13] = g
        return full, None, None


def central_differences(prob, theta, period, h_t=1e-4, h_r=1e-6, h_d=1e-10, facets=None,
                        precision=64):
    """Central finite differences of L = mu + sigma for each facet parameter
    (columns: dx, dy, dz, rx, ry, rz, period), with the photon paths frozen.

    Each difference is formed from per-photon coordinate differences, not as
    the difference of two loss values: moving facet f changes only its own
    photons' coordinates, by Delta_i, and then exactly
        mu(+-) - mu0 = d = sum_f w Delta / W,
        sigma(+-)^2 - sigma0^2 = -d^2 + sum_f w Delta (2 (y0 - mu0) + Delta) / W,
    so the result keeps the digits of Delta (a loss of ~500 mm differenced
    over a step of 1e-4 mm would lose nine of them).  Only facet f's photons
    are replayed for its steps."""
    F = theta.shape[0]
    facets = range(F) if facets is None else facets
    out = torch.full((F, 7), float('nan'), dtype=torch.float64, device=theta.device)
    steps = [h_t] * 3 + [h_r] * 3 + [h_d]
    with torch.no_grad():
        g0 = prob.geometry(theta, period).contiguous()
        y0 = prob.coordinates(g0, precision)
        ok = ~torch.isnan(y0) & (prob.weight != 0)
        w = torch.where(ok, prob.weight, torch.zeros_like(prob.weight))
        W = w.sum()
        mu0 = (w * torch.where(ok, y0, torch.zeros_like(y0))).sum() / W
        var0 = (w * torch.where(ok, y0 - mu0, torch.zeros_like(y0)) ** 2).sum() / W
        for f in facets:
            idx = torch.nonzero(ok & (prob.facet == f)).flatten()
            if idx.numel() == 0:
                out[f] = 0.
                continue
            src = {'n': int(idx.numel()), 'pos': prob.pos[:, idx].contiguous(),
                   'dir': prob.dir[:, idx].contiguous(), 'energy': prob.energy[idx].contiguous(),
                   'weight': w[idx].contiguous(), 'facet': prob.facet[idx].contiguous(),
                   'order': prob.order[idx].contiguous(), 'ccd': prob.ccd[idx].contiguous()}
            wf, yf = w[idx], y0[idx]
            for k in range(7):
                dmu, dvar = [], []
                for sgn in (1., -1.):
                    th = theta.clone()
                    pe = period.clone()
                    if k < 6:
                        th[f, k] += sgn * steps[k]
                    else:
                        pe[f] += sgn * steps[k]
                    g = prob.geometry(th, pe).contiguous()
                    y = torch.empty(src['n'], dtype=torch.float64, device=prob.device)
                    _launch(prob._args(g, precision, y_out=y, src=src),
                            torch.cuda.current_stream(prob.device).cuda_stream)
                    delta = y - yf
                    d = (wf * delta).sum() / W
                    dmu.append(d)
                    dvar.append(-d * d + (wf * delta * (2 * (yf - mu0) + delta)).sum() / W)
                sp = torch.sqrt(var0 + dvar[0])
                sm = torch.sqrt(var0 + dvar[1])
                dL = (dmu[0] - dmu[1]) + (dvar[0] - dvar[1]) / (sp + sm)
                out[f, k] = dL / (2 * steps[k])
    return out


def entrance_chunk(ci, front, src, uniforms, weight_fn, precision=64):
    """Adjoint inputs for one chunk of photons, traced on demand: ``ci``
    traces the whole instrument (frozen facet, order and CCD choices, the
    traced coordinate), ``front`` the elements before the grating array
    (the photons as they reach it), both from the same ``uniforms``.
    ``weight_fn(traced DevicePhotons)`` gives each photon's weight in the
    loss (0 excludes it).  Returns the dict :class:`ChunkedAdjoint` takes
    (``y_traced``: all three traced position components)."""
    from .tracer import DevicePhotons
    full = DevicePhotons(src.dir, src.pol, src.energy, src.prob, device=src.device)
    ci.trace_device(full, uniforms=uniforms, precision=precision)
    ent = DevicePhotons(src.dir, src.pol, src.energy, src.prob, device=src.device)
    front.trace_device(ent, uniforms=uniforms[:front.n_uniform_slots].contiguous(), precision=precision)
    return {'pos': ent.pos, 'dir': ent.dir, 'energy': ent.energy, 'weight': weight_fn(full),
            'facet': full.icols['facet'], 'order': full.icols['order'], 'ccd': full.icols['CCD'],
            'y_traced': full.pos}


class ChunkedAdjoint:
    """L = centroid + RMS width and its gradient over more photons than are
    held at once: the photons come in chunks that ``chunk(k)`` recomputes on
    demand (e.g. :func:`entrance_chunk` tracing chunk k's photons) and that
    are replayed one after another into the same per-facet sums A, B and
    moments; the pose chain rule runs once at the end.  Only one chunk's
    entrance state (76 bytes per photon) lives at a time.  The sums are
    those of the one-pass :meth:`AdjointProblem.launch_loss_and_grad`, so
    the result equals it over the union of the chunks up to the order of
    the additions.

    ``chunk(k)`` returns a dict of device tensors pos, dir (3, n), energy,
    weight (n), facet, order, ccd (int32, n) and optionally ``y_traced``
    ((3, n) traced positions: photons whose one-facet replay does not
    reproduce them are excluded, as :meth:`AdjointProblem.restrict_to_traced`
    does).  The moments are shifted by the weighted mean coordinate of chunk
    0 at the unperturbed pose.
    """

    def __init__(self, gratings, detectors, chunk, n_chunks, coord=1, device=None):
        self.chunk, self.n_chunks = chunk, int(n_chunks)
        c0 = chunk(0)
        self.prob = AdjointProblem(gratings, detectors, c0['pos'], c0['dir'], c0['energy'],
                                   c0['weight'], c0['facet'], c0['order'], c0['ccd'], coord=coord,
                                   device=device)
        if 'y_traced' in c0:
            self.prob.restrict_to_traced(c0['y_traced'][coord])
        self.prob._pose_buffers()
        self.geom0 = self.prob.geometry(self.prob.theta0, self.prob.period0).contiguous()
        self.n_used = 0

    def _compact(self, c):
        p = self.prob
        f64 = dict(dtype=torch.float64, device=p.device)
        pos = c['pos'].to(**f64).contiguous()
        dir_ = c['dir'].to(**f64).contiguous()
        energy = c['energy'].to(**f64).contiguous()
        weight = c['weight'].to(**f64).contiguous()
        i32 = dict(dtype=torch.int32, device=p.device)
        facet, order, ccd = (c[k].to(**i32).contiguous() for k in ('facet', 'order', 'ccd'))
        use = (weight != 0) & (facet >= 0) & (ccd >= 0)
        if 'y_traced' in c:
            # the unperturbed one-facet replay must reproduce the trace
            y = torch.empty(pos.shape[1], dtype=torch.float64, device=p.device)
            whole = {'n': pos.shape[1], 'pos': pos, 'dir': dir_, 'energy': energy, 'weight': weight,
                     'facet': facet, 'order': order, 'ccd': ccd}
            _launch(p._args(self.geom0, 64, y_out=y, src=whole),
                    torch.cuda.current_stream(p.device).cuda_stream)
            yt = c['y_traced'][p.coord].to(**f64)
            use = use & (torch.abs(y - yt) <= 1e-10 * torch.abs(yt).clamp(min=1.0))
        idx = torch.nonzero(use).flatten()
        return {'n': int(idx.numel()), 'pos': pos[:, idx].contiguous(), 'dir': dir_[:, idx].contiguous(),
                'energy': energy[idx].contiguous(), 'weight': weight[idx].contiguous(),
                'facet': facet[idx].contiguous(), 'order': order[idx].contiguous(),
                'ccd': ccd[idx].contiguous()}

    def loss_and_grad(self, theta, period, precision=32):
        """(L, dL/dtheta (F, 6), dL/dperiod (F,)) over all chunks."""
        p = self.prob
        pb = p._pose_buffers()
        stream = torch.cuda.current_stream(p.device).cuda_stream
        pb['theta'].copy_(theta)
        pb['period'].copy_(period)
        pb['accum'].zero_()
        _lib.check(_lib.lib.mx_facet_pose(ctypes.byref(p._pose_args(0)), ctypes.c_void_p(stream)))
        self.n_used = 0
        for k in range(self.n_chunks):
            src = self._compact(self.chunk(k))
            self.n_used += src['n']
            if src['n'] == 0:
                continue
            one = p._args(pb['geom'], precision, grad=pb['accum'], moments=pb['moments'], src=src)
            one.flags = _lib.MX_AJ_SUMS
            one.lds_accumulate = int(p.F <= 315)
            one.y_shift = p.y_shift
            _launch(one, stream)
            del src
        _lib.check(_lib.lib.mx_facet_pose(ctypes.byref(p._pose_args(2)), ctypes.c_void_p(stream)))
        return pb['loss'][0].clone(), pb['grad7'][:, :6].clone(), pb['grad7'][:, 6].clone()
