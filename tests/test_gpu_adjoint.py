"""Adjoint of the detector line centroid + width w.r.t. facet poses and
periods (needs an MI355X), on the differentiation workload: 200-facet
Rowland array with N(0, 0.1 mm) / N(0, 30 arcsec) facet perturbations,
1e6 photons at 0.65 keV, loss = centroid + RMS width of order -6.

Checks: the replay coordinate equals the parity-tested trace; the FP64
adjoint agrees entrywise with central finite differences (frozen paths) to
<= 1e-4 relative on every facet parameter whose gradient exceeds 1e-3 of
its column's largest; the FP32 adjoint agrees with the FP64 adjoint to
within the relative error its FP32 detector coordinates put into the seeds
(rms(y32 - y64) / sigma, at least 1e-4); the chunked replay (photons
traced chunk by chunk, never held together) equals the one-shot gradient."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def problem(gpu):
    import torch
    from marxs_amd import configs, tolerancing
    from marxs_amd.adjoint import AdjointProblem
    from marxs_amd.tracer import CompiledInstrument, DevicePhotons
    import marxs_amd as mx
    inst, parts = configs.cat_rowland_instrument(200)
    tolerancing.wiggle(parts['gratings'], dx=0.1, dy=0.1, dz=0.1, rx=1.4544e-4, ry=1.4544e-4,
                       rz=1.4544e-4, rng=np.random.default_rng(5))
    n = 1_000_000
    rng = np.random.default_rng(6)
    ci = CompiledInstrument(inst)
    u = rng.random((ci.n_uniform_slots + 1, n))
    ph = configs.onaxis_photons(np.full(n, 0.65), u[-1])
    uni = torch.as_tensor(np.ascontiguousarray(u[:ci.n_uniform_slots]), device=gpu)
    full = DevicePhotons(ph['dir'].T, ph['pol'].T, ph['energy'], ph['prob'], device=gpu)
    ci.trace_device(full, uniforms=uni)
    front = CompiledInstrument(mx.Sequence(elements=[parts['aperture'], parts['lens']]))
    ent = DevicePhotons(ph['dir'].T, ph['pol'].T, ph['energy'], ph['prob'], device=gpu)
    front.trace_device(ent, uniforms=uni[:2].contiguous())
    sel = (full.icols['order'] == -6) & (full.icols['CCD'] >= 0)
    weight = torch.where(sel, full.prob, torch.zeros_like(full.prob))
    prob = AdjointProblem(parts['gratings'], parts['ccds'], ent.pos, ent.dir, ent.energy, weight,
                          full.icols['facet'], full.icols['order'], full.icols['CCD'], coord=1)
    y = prob.coordinates(prob.geometry(prob.theta0, prob.period0))
    dev = (torch.abs(y - full.pos[1]) / torch.abs(full.pos[1]))[sel]
    print('replay vs trace: max rel', float(dev.max()), 'n >1e-10', int((dev > 1e-10).sum()),
          'of', int(sel.sum()))
    prob.n_selected = int(sel.sum())
    prob.n_dropped = prob.restrict_to_traced(full.pos[1])
    sel = sel & (prob.weight != 0)
    prob.inputs = (inst, parts, ph, uni)
    return prob, full, sel


def test_replay_matches_trace(problem):
    prob, full, sel = problem
    import torch
    y = prob.coordinates(prob.geometry(prob.theta0, prob.period0))
    ok = sel.cpu().numpy()
    assert ok.sum() > 1000
    # The one-facet replay reproduces the traced coordinate to 1e-10 except
    # for photons that met two facets (or two CCDs) in the trace; those few
    # are excluded from the loss.
    assert prob.n_dropped <= 0.01 * prob.n_selected
    yt = full.pos[1].cpu().numpy()
    yr = y.cpu().numpy()
    assert np.allclose(yr[ok], yt[ok], rtol=1e-10, atol=1e-9)
    assert np.isnan(yr[~ok]).all()


def test_adjoint_vs_finite_differences(problem):
    """Every facet parameter whose gradient is at least 1e-3 of its column's
    largest: |adjoint - central difference| <= 1e-4 |central difference|."""
    import torch
    prob, full, sel = problem
    th = prob.theta0.clone().requires_grad_(True)
    pe = prob.period0.clone().requires_grad_(True)
    L = prob.loss(th, pe, precision=64)
    L.backward()
    g = torch.cat([th.grad, pe.grad[:, None]], dim=1)
    from marxs_amd.adjoint import central_differences
    fd = central_differences(prob, prob.theta0, prob.period0)
    n_checked = 0
    for k, name in enumerate(['dx', 'dy', 'dz', 'rx', 'ry', 'rz', 'period']):
        big = torch.abs(g[:, k]) >= 1e-3 * torch.abs(g[:, k]).max()
        a, b = g[big, k], fd[big, k]
        err = torch.abs(a - b) / torch.abs(b)
        n_checked += int(big.sum())
        print(name, 'checked', int(big.sum()), 'max rel', float(err.max()))
        assert float(err.max()) <= 1e-4, (name, float(err.max()))
    assert n_checked >= 7 * 150


def test_fp32_adjoint_matches_fp64(problem):
    import torch
    prob, full, sel = problem
    grads = {}
    for prec in (64, 32):
        th = prob.theta0.clone().requires_grad_(True)
        pe = prob.period0.clone().requires_grad_(True)
        prob.loss(th, pe, precision=prec).backward()
        grads[prec] = torch.cat([th.grad, pe.grad[:, None]], dim=1)
    g64 = prob.geometry(prob.theta0, prob.period0)
    y64 = prob.coordinates(g64, 64)
    y32 = prob.coordinates(g64, 32)
    ok = ~torch.isnan(y64)
    w = prob.weight[ok]
    mu = (w * y64[ok]).sum() / w.sum()
    sigma = float(torch.sqrt((w * (y64[ok] - mu) ** 2).sum() / w.sum()))
    rms = float(torch.sqrt(((y32 - y64)[ok] ** 2).mean()))
    print('sigma', sigma, 'fp32 y rms error', rms, 'max', float(torch.abs(y32 - y64)[ok].max()))
    # The seeds dL/dy = w/W (1 + (y - mu)/sigma) inherit the FP32 coordinate
    # error: relative error ~ rms(y32 - y64)/sigma (~1e-3 for a 60 um line and
    # 0.07 um FP32 error), which bounds the FP32 gradient error.
    tol = max(1e-4, rms / sigma)
    for k in range(7):
        scale = torch.abs(grads[64][:, k]).max()
        err = float((torch.abs(grads[32][:, k] - grads[64][:, k]) / scale).max())
        print(k, float(scale), err, tol)
        assert err <= tol, (k, err, tol)


def _autograd(prob, th0, pe0, prec):
    import torch
    th = th0.clone().requires_grad_(True)
    pe = pe0.clone().requires_grad_(True)
    L = prob.loss(th, pe, precision=prec)
    L.backward()
    return L.detach(), th.grad, pe.grad


def test_fused_loss_and_grad_match_autograd(problem):
    """The fused path (pose kernel, forward replay with moments, adjoint with
    seeds from the moments, pose chain rule in HIP) equals the autograd path
    in FP64, and its FP32 version stays within the FP32 seed tolerance."""
    import torch
    prob, full, sel = problem
    rng = torch.Generator(device=prob.device).manual_seed(3)
    th = prob.theta0 + 1e-3 * torch.randn(prob.theta0.shape, generator=rng, dtype=torch.float64,
                                          device=prob.device) * torch.tensor([50., 50, 50, .1, .1, .1],
                                                                            dtype=torch.float64,
                                                                            device=prob.device)
    pe = prob.period0.clone()
    La, gta, gpa = _autograd(prob, th, pe, 64)
    Lf, gtf, gpf = prob.loss_and_grad(th, pe, precision=64)
    assert abs(float(Lf - La)) <= 1e-10 * abs(float(La))
    for a, b in ((gta, gtf), (gpa[:, None], gpf[:, None])):
        for k in range(a.shape[1]):
            scale = float(torch.abs(a[:, k]).max())
            assert float(torch.abs(a[:, k] - b[:, k]).max()) <= 1e-9 * scale, k
    L32, gt32, gp32 = prob.loss_and_grad(th, pe, precision=32)
    assert abs(float(L32 - La)) <= 1e-6 * abs(float(La))
    for a, b in ((gta, gt32), (gpa[:, None], gp32[:, None])):
        for k in range(a.shape[1]):
            scale = float(torch.abs(a[:, k]).max())
            assert float(torch.abs(a[:, k] - b[:, k]).max()) <= 2e-3 * scale, k


def test_fused_graph_replays_follow_the_parameters(problem):
    import torch
    prob, full, sel = problem
    g = prob.capture(precision=64)
    buf = prob.buffers
    for rep in range(3):
        th = prob.theta0.clone()
        th[:, 0] += 0.05 * rep
        th[:, 5] += 1e-4 * rep
        buf['theta'].copy_(th)
        buf['period'].copy_(prob.period0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = (buf['loss'][0].item(), buf['grad7'].clone())
        # eager evaluation in between replays
        L, gt, gp = prob.loss_and_grad(th, prob.period0, precision=64)
        # the loss comes from moments summed with atomic adds: equal up to
        # the summation order
        assert abs(got[0] - L.item()) <= 1e-13 * abs(L.item())
        # gradients are sums of atomic adds: equal up to summation order
        ref = torch.cat([gt, gp[:, None]], 1)
        scale = torch.abs(ref).max(0).values
        assert bool((torch.abs(got[1] - ref) <= 1e-12 * scale).all())
        if rep:
            assert not torch.equal(got[1], prev)
        prev = got[1]


def test_chunked_replay_equals_one_shot(problem):
    """The same 1e6 photons traced and replayed in four chunks of 2.5e5
    (ChunkedAdjoint over entrance_chunk: each chunk's entrance state and
    frozen choices recomputed by a trace, then accumulated) give the one-shot
    loss and gradient (the sums differ only in the order of the additions)."""
    import torch
    import marxs_amd as mx
    from marxs_amd.adjoint import ChunkedAdjoint, entrance_chunk
    from marxs_amd.tracer import CompiledInstrument, DevicePhotons
    prob, full, sel = problem
    inst, parts, ph, uni = prob.inputs
    ci = CompiledInstrument(inst)
    front = CompiledInstrument(mx.Sequence(elements=[parts['aperture'], parts['lens']]))
    n, m = len(ph['energy']), 250_000

    def chunk(k):
        s = slice(k * m, (k + 1) * m)
        src = DevicePhotons(ph['dir'][s].T, ph['pol'][s].T, ph['energy'][s], ph['prob'][s],
                            device=prob.device)

        def weight(t):
            return torch.where((t.icols['order'] == -6) & (t.icols['CCD'] >= 0), t.prob,
                               torch.zeros_like(t.prob))
        return entrance_chunk(ci, front, src, uni[:, s].contiguous(), weight)

    ca = ChunkedAdjoint(parts['gratings'], parts['ccds'], chunk, n // m, coord=1)
    rng = torch.Generator(device=prob.device).manual_seed(4)
    th = prob.theta0 + 1e-4 * torch.randn(prob.theta0.shape, generator=rng, dtype=torch.float64,
                                          device=prob.device)
    for prec in (64, 32):
        Lc, gtc, gpc = ca.loss_and_grad(th, prob.period0, precision=prec)
        Lo, gto, gpo = prob.loss_and_grad(th, prob.period0, precision=prec)
        assert ca.n_used == int((prob.weight != 0).sum())
        assert abs(float(Lc - Lo)) <= 1e-12 * abs(float(Lo))
        for a, b in ((gto, gtc), (gpo[:, None], gpc[:, None])):
            for k in range(a.shape[1]):
                scale = float(torch.abs(a[:, k]).max())
                assert float(torch.abs(a[:, k] - b[:, k]).max()) <= 1e-10 * scale, (prec, k)
