#!/usr/bin/env python
"""Benchmark: photons/s through a 600-facet CAT array + 16 CCDs on one MI355X.

Workload (one step = one trace of the whole batch): 1e7 photons, FixedPointing
on axis, line spectrum (0.5, 0.65, 0.9, 1.2 keV equally weighted plus a 20 %
flat 0.3-2 keV continuum), annulus aperture 300-600 mm and PerfectLens
(f = 12500 mm), 600 CAT facets (25 x 25 mm, d = 2e-4 mm, blaze 1.5 deg) on a
Rowland torus (R = 5900, r = 6000 mm), orders -12..+2 from a synthetic
Gaussian efficiency table, 16 CCDs of 1024^2 x 0.024 mm on the Rowland
circle.  Inputs are generated on the GPU and resident in HBM before the timed
region; every step reads the same input columns and writes fresh outputs.

``value`` is the FP32 local-frame tracer's photons/s; the FP64 parity
kernel's rate, their ratio, the FP32 detector error against FP64 and the CPU
oracle's rate on a bounded sample are reported beside it, and
``fp32_with_draws`` times the FP32 step with its uniform draws generated on
the device inside the timed region.

Two more sections run after it (``--sections``):

``adjoint``  200 CAT facets, 1e6 photons at 0.65 keV, facet poses perturbed by
             N(0, 0.1 mm) and N(0, 30 arcsec); one step = loss (centroid + RMS
             width of order -6 on the CCDs) and its gradient with respect to
             200 x 6 pose parameters and 200 periods (FP32 adjoint kernel,
             torch autograd over the pose chain).  Reports adjoints/s, the
             agreement with central finite differences and with the FP64
             adjoint.
``stack``    the trace workload with every facet a CATL1L2Stack (membrane,
             quality factor, L1 cross-support with the Si bar transmission,
             L2 absorption and L2 diffraction), 1e7 photons: FP32 and FP64
             rates (FP64 as the split trace, the default there), the other
             trace form of each precision, FP64 indices against the oracle on
             a sample, the oracle's rate on that sample.
``interp``   the trace workload with a blaze-dependent
             InterpolateEfficiencyTable (wavelength x blaze angle, orders
             -12..2) as every facet's order selector: FP32 and FP64 rates,
             FP64 indices against the oracle on a sample.
``stages``   (not in the default list) the trace workload cut after each
             stage (aperture; + lens; + facets; + CCDs without facets; all),
             FP32 and FP64 kernel times: where the trace time goes.
``pipeline`` a full analysis run: aperture, lens, RadialMirrorScatter at the
             lens (2e-5 / 4e-6 rad), the 600 CAT facets, catsupportbars and a
             CircularDetector on the Rowland circle, 1e7 photons at 0.65 keV
             (resolving power is a per-energy figure); FP32 / FP64
             trace rates, then CaptureResAeff (effective area and
             sigma-clipped resolving power of orders -12..2) on the device
             columns, against the numpy oracle of the figures of merit on the
             same columns.
``batched``  64 design variants (Rowland R ~ U(5000, 7000) mm, blaze ~
             U(1, 2) deg, 50-800 facets, own CCD rows), 2e5 photons each at
             0.3-2.0 keV, traced in one autotuned, hipGraph-replayed launch,
             against the same variants launched one after the other; the FP64
             batched results are compared with the sequential FP64 traces.

    python bench.py [--steps K] [--warmup W] [--photons N] [--cpu-sample S]
                    [--sections trace,adjoint,stack,interp,pipeline,batched]
"""
import argparse
import json
import os
import re
import sys
import time

HBM_PEAK_GBS = 8000.0
FP32_VALU_PEAK_TFLOPS = 157.3
FP64_VALU_PEAK_TFLOPS = 78.6
FLOPS_PER_TEST = 33   # ray-plane test: (c-p).n 8, d.n 5, divide 1, p+k d 6, q-c 3, two dots 10
METRIC = 'photons/sec, 600-facet CAT array: FP32 vs FP64 GPU at stated µm error; adjoint/s'

# algorithmic bytes per photon, FP64 columns, each touched once:
# read  dir 24, polarization 24, energy 8, probability 8, uniforms 3 x 8
# write pos 24, dir 24, polarization 24, probability 8, y/z 16, mirror_x/y 16,
#       facet 4, order 4, blaze 8, grat_y/z 16, CCD 4, det_x/y 16,
#       detpix_x/y 16, pix_x/y 8
READ_BYTES = 24 + 24 + 8 + 8 + 3 * 8
WRITE_BYTES = 24 + 24 + 24 + 8 + 16 + 16 + 4 + 4 + 8 + 16 + 4 + 16 + 16 + 8
BYTES_PER_PHOTON = READ_BYTES + WRITE_BYTES


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--photons', type=int, default=10_000_000)
    ap.add_argument('--facets', type=int, default=600)
    ap.add_argument('--cpu-sample', type=int, default=1_000_000)
    ap.add_argument('--seed', type=int, default=20261202)
    ap.add_argument('--grid', type=int, default=1)
    ap.add_argument('--only', choices=['fp32', 'fp64'], default=None,
                    help='profiling: time one precision only, skip comparisons')
    ap.add_argument('--sections', default='trace,adjoint,stack,interp,pipeline,batched')
    ap.add_argument('--stack-cpu-sample', type=int, default=100_000)
    ap.add_argument('--cpu-procs', type=int, default=0,
                    help='processes for the multi-process CPU baseline (0 = min(16, cpu count))')
    ap.add_argument('--variants', type=int, default=64)
    ap.add_argument('--variant-photons', type=int, default=200_000)
    return ap.parse_args()


def _oracle_shard(job):
    """One process of the multi-process CPU baseline: trace a photon shard
    with the numpy oracle (imports no GPU library)."""
    els, ph, us = job
    os.environ['OMP_NUM_THREADS'] = '1'
    from oracle import trace as OT
    t = time.perf_counter()
    OT.trace(els, ph, us)
    return time.perf_counter() - t


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_profile_value(root, name, key):
    """A number committed in a profiles/ summary as 'key: X' (figures that
    come from rocprofv3 counter runs of this bench: HBM traffic per launch,
    VALU instruction counts), or None."""
    path = os.path.join(root, 'profiles', name)
    if not os.path.exists(path):
        return None
    m = re.search(re.escape(key) + r':\s*([0-9.eE+-]+)', open(path).read())
    return float(m.group(1)) if m else None


def read_traffic(root):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by a copy-kernel calibration), committed in
    profiles/trace_c3_fp32_pmc.txt as 'traffic_bytes_per_launch: X'."""
    return read_profile_value(root, 'trace_c3_fp32_pmc.txt', 'traffic_bytes_per_launch')


def valu_from_counters(root, prec, kernel_s, n):
    """VALU figures of the headline kernel from the committed counter run
    (profiles/trace_c3_valu_pmc.txt: per-dispatch SQ_INSTS_VALU and the
    FP32/FP64 add, multiply, FMA and transcendental instruction counts of
    the same bench command): the fraction of the 1024 SIMDs' issue cycles
    that VALU instructions take (4 cycles per wave64 instruction at
    2.4 GHz, over this run's kernel time) and the achieved FLOP/s against
    the FP32 / FP64 vector peak (64 lanes per instruction, an FMA counts 2)."""
    tag = 'fp32' if prec == 32 else 'fp64'
    insts = read_profile_value(root, 'trace_c3_valu_pmc.txt', f'valu_insts_per_dispatch_{tag}')
    flop = read_profile_value(root, 'trace_c3_valu_pmc.txt', f'valu_flops_per_dispatch_{tag}')
    if insts is None or flop is None:
        return None
    peak = FP32_VALU_PEAK_TFLOPS if prec == 32 else FP64_VALU_PEAK_TFLOPS
    return {'valu_insts_per_launch': insts,
            'valu_insts_per_photon_wave': insts / (n / 64),
            'valu_issue_fraction': insts * 4 / (1024 * 2.4e9 * kernel_s),
            'flops_per_launch': flop,
            'tflops': flop / kernel_s / 1e12,
            'flop_fraction_of_peak': flop / kernel_s / 1e12 / peak,
            'peak_tflops': peak,
            'source': 'profiles/trace_c3_valu_pmc.txt (rocprofv3 --pmc of this bench command)'}


def main():
    args = parse()
    if args.gpus != 1:
        print(f'bench.py runs on exactly one GPU (asked for {args.gpus}); the tracer '
              'shards by photon ranges across GPUs but no multi-GPU run is part of this build.',
              file=sys.stderr)
        sys.exit(2)
    root = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, root)
    import numpy as np
    import torch
    from marxs_amd import configs
    from marxs_amd.tracer import CompiledInstrument, DevicePhotons, raise_for_status

    dev = torch.device('cuda', 0)
    torch.cuda.set_device(dev)
    n = args.photons
    t0 = time.time()
    inst, parts = configs.cat_rowland_instrument(args.facets)
    ci = CompiledInstrument(inst, grid=bool(args.grid), device=dev)
    n_elem = len(ci.tables.elements) - 1   # elements tested per photon (not the aperture)
    log(f'instrument: {len(parts["gratings"].elements)} facets, {len(parts["ccds"].elements)} CCDs, '
        f'compiled in {time.time() - t0:.2f}s')

    # ---- synthetic inputs, generated and resident on the GPU
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed)
    ns = ci.n_uniform_slots
    u = torch.rand((ns + 3, n), dtype=torch.float64, device=dev, generator=g)
    lines = torch.tensor(configs.LINES, dtype=torch.float64, device=dev)
    cont = u[0] < 0.2
    idx = torch.clamp(((u[0] - 0.2) / 0.2).long(), 0, 3)
    energy = torch.where(cont, 0.3 + 1.7 * u[1], lines[idx])
    phi = 2 * np.pi * u[2]
    dirs = torch.zeros((3, n), dtype=torch.float64, device=dev)
    dirs[0] = -1.
    pol = torch.stack([torch.zeros_like(phi), torch.sin(phi), torch.cos(phi)])
    uni = u[3:].contiguous()
    src = DevicePhotons(dirs, pol, energy, torch.ones(n, dtype=torch.float64, device=dev),
                        device=dev)
    del dirs, pol, phi, idx, cont
    out32 = DevicePhotons(src.dir, src.pol, src.energy, src.prob, device=dev)
    out64 = DevicePhotons(src.dir, src.pol, src.energy, src.prob, device=dev)
    status = torch.zeros(8, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()

    # the persistent grid (workgroups per CU) measured on this GPU before the
    # timed steps (marxs_amd.autotune.tune_trace: 1 to 4 per CU, median of 5
    # launches each); every later launch of this instrument uses it
    from marxs_amd.autotune import tune_trace
    grid_tuned = {}
    for prec, out in ((32, out32), (64, out64)):
        if args.only and args.only != f'fp{prec}':
            continue
        best, table = tune_trace(ci, out, uni, prec, src=src)
        grid_tuned[f'fp{prec}'] = {'wg_per_cu': best, 'ms': {str(k): round(v, 4) for k, v in table.items()}}
    log('grid tuned', grid_tuned)

    def step(out, prec):
        ci.launch(out, uni, precision=prec, src=src, status=status)

    def timed(out, prec, fn=step):
        for _ in range(args.warmup):
            fn(out, prec)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
        t = time.perf_counter()
        for a, b in ev:
            a.record()
            fn(out, prec)
            b.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        kern = [a.elapsed_time(b) * 1e-3 for a, b in ev]
        return wall, float(np.mean(kern))

    sections = set(args.sections.split(','))
    if args.only:
        prec = 32 if args.only == 'fp32' else 64
        wall, k = timed(out32 if prec == 32 else out64, prec)
        raise_for_status(status.cpu().numpy())
        log(f'{args.only}: kernel {k * 1e3:.3f} ms, {n / k:.4e} photons/s')
        return
    wall32, k32 = timed(out32, 32)
    log(f'FP32: {wall32 / args.steps * 1e3:.3f} ms/step, kernel {k32 * 1e3:.3f} ms')
    wall64, k64 = timed(out64, 64)
    log(f'FP64: {wall64 / args.steps * 1e3:.3f} ms/step, kernel {k64 * 1e3:.3f} ms')
    # the same FP32 step with its uniform draws generated on the device
    # inside the timed region (the reference draws its random numbers while
    # processing), into buffers of their own
    uni_d = torch.empty_like(uni)
    out_d = DevicePhotons(src.dir, src.pol, src.energy, src.prob, device=dev)
    status_d = torch.zeros_like(status)
    gd = torch.Generator(device=dev)
    gd.manual_seed(args.seed + 1)

    def step_draws(out, prec):
        uni_d.uniform_(0.0, 1.0, generator=gd)
        ci.launch(out, uni_d, precision=prec, src=src, status=status_d)
    wall_d, _ = timed(out_d, 32, step_draws)
    raise_for_status(status_d.cpu().numpy())
    del out_d, uni_d
    log(f'FP32 with draws generated in the step: {wall_d / args.steps * 1e3:.3f} ms/step')
    st = status.cpu().numpy()
    raise_for_status(st)

    # ---- FP32 error against the FP64 GPU trace (whole batch)
    same = ((out32.icols['CCD'] == out64.icols['CCD']) & (out64.icols['CCD'] >= 0)
            & (out32.icols['facet'] == out64.icols['facet']))
    dx = (out32.fcols['det_x'] - out64.fcols['det_x'])[same]
    dy = (out32.fcols['det_y'] - out64.fcols['det_y'])[same]
    e_um = torch.sqrt(dx * dx + dy * dy) * 1e3
    dd = (out32.dir - out64.dir)[:, same]
    ang = torch.sqrt((dd * dd).sum(0)) * 206264.806
    hit64 = out64.icols['CCD'] >= 0
    err = {
        'rms_um': float(torch.sqrt(torch.mean(e_um ** 2))),
        'max_um': float(e_um.max()),
        'dir_rms_arcsec': float(torch.sqrt(torch.mean(ang ** 2))),
        'dir_max_arcsec': float(ang.max()),
        'facet_edge_mismatch': int((out32.icols['facet'] != out64.icols['facet']).sum()),
        'ccd_edge_mismatch': int((out32.icols['CCD'] != out64.icols['CCD']).sum()),
        'pixel_mismatch': int(((out32.icols['pix_x'] != out64.icols['pix_x']) |
                               (out32.icols['pix_y'] != out64.icols['pix_y']))[same].sum()),
        'compared_photons': int(same.sum()),
        'photons_on_ccds': int(hit64.sum()),
        'reference': 'FP64 GPU kernel (bit-level parity with the oracle)',
    }
    log('FP32 error', err)

    # ---- CPU baseline: the numpy oracle on a bounded sample
    cpu = None
    if args.cpu_sample > 0:
        os.environ.setdefault('OMP_NUM_THREADS', '1')
        from oracle import trace as OT
        from oracle.describe import describe
        s = min(args.cpu_sample, n)
        ph = {'dir': src.dir[:, :s].cpu().numpy().T.copy(), 'pol': src.pol[:, :s].cpu().numpy().T.copy(),
              'energy': src.energy[:s].cpu().numpy(), 'prob': src.prob[:s].cpu().numpy()}
        us = uni[:, :s].cpu().numpy()
        els = describe(inst)
        t = time.perf_counter()
        ref = OT.trace(els, ph, us)
        tc = time.perf_counter() - t
        ok = (ref['CCD'] >= 0) & (ref['CCD'] == out32.icols['CCD'][:s].cpu().numpy())
        ddx = (out32.fcols['det_x'][:s].cpu().numpy() - ref['det_x'])[ok]
        ddy = (out32.fcols['det_y'][:s].cpu().numpy() - ref['det_y'])[ok]
        err['vs_oracle_sample_rms_um'] = float(np.sqrt(np.mean(ddx ** 2 + ddy ** 2)) * 1e3)
        f64ok = np.array_equal(ref['CCD'], out64.icols['CCD'][:s].cpu().numpy()) and \
            np.array_equal(ref['facet'], out64.icols['facet'][:s].cpu().numpy()) and \
            np.array_equal(ref['order'], out64.icols['order'][:s].cpu().numpy().astype(np.int64))
        err['fp64_sample_indices_equal_oracle'] = bool(f64ok)
        cpu = {'value': s / tc, 'unit': 'photons/s', 'cores': 1, 'kind': 'port',
               'sample': f'first {s} photons of the same batch (same instrument, energies, '
                         f'uniforms), numpy float64 element-major oracle, one thread, '
                         f'{tc:.1f} s'}
        log('CPU oracle', cpu)
        # the reference's EfficiencyFile selects orders one photon at a time
        # (grating.py:90-96): the same oracle with that loop, on a tenth of
        # the sample, identical results
        s_loop = max(1, s // 10)
        ph_l = {k: v[:s_loop] for k, v in ph.items()}
        t = time.perf_counter()
        ref_l = OT.trace(els, ph_l, us[:, :s_loop], selector_loop=True)
        tl = time.perf_counter() - t
        same = all(np.array_equal(np.asarray(ref_l[k]), np.asarray(ref[k])[:s_loop],
                                  equal_nan=np.asarray(ref_l[k]).dtype.kind == 'f') for k in ref_l)
        cpu_loop = {'value': s_loop / tl, 'unit': 'photons/s', 'cores': 1, 'kind': 'port',
                    'sample': f'first {s_loop} photons, the oracle with the per-photon order '
                              f'selection loop of the reference\'s EfficiencyFile, one thread, '
                              f'{tl:.1f} s', 'identical_to_vectorised': bool(same)}
        log('CPU oracle, per-photon selector loop', cpu_loop)
        # the same sample split over processes (photons are independent)
        import multiprocessing as mp
        procs = args.cpu_procs or min(16, os.cpu_count() or 1)
        edges = np.linspace(0, s, procs + 1).astype(int)
        jobs = [(els, {k: v[a:b] for k, v in ph.items()}, us[:, a:b])
                for a, b in zip(edges[:-1], edges[1:])]
        t = time.perf_counter()
        with mp.get_context('spawn').Pool(procs) as pool:
            pool.map(_oracle_shard, jobs)
        tp = time.perf_counter() - t
        cpu_mp = {'value': s / tp, 'unit': 'photons/s', 'cores': procs, 'kind': 'port',
                  'sample': f'the same {s} photons in {procs} spawned processes '
                            f'(one contiguous shard each, one thread each), {tp:.1f} s '
                            'including process start-up', 'cpu_count': os.cpu_count()}
        log('CPU oracle, processes', cpu_mp)

    # ---- the same trace through the host photon-table API (PCIe included)
    host_rate = None
    if 'hosttable' in sections or 'trace' in sections:
        import marxs_amd as mx
        m = min(n, 1_000_000)
        tab = mx.PhotonTable({'dir': np.c_[src.dir[:, :m].cpu().numpy().T, np.zeros(m)],
                              'polarization': np.c_[src.pol[:, :m].cpu().numpy().T, np.zeros(m)],
                              'energy': src.energy[:m].cpu().numpy(),
                              'probability': src.prob[:m].cpu().numpy()})
        um = uni[:, :m].cpu().numpy()
        ci.process_photons(tab.copy(), uniforms=um, precision=32)
        torch.cuda.synchronize()
        t = time.perf_counter()
        ci.process_photons(tab.copy(), uniforms=um, precision=32)
        torch.cuda.synchronize()
        th = time.perf_counter() - t
        host_rate = {'photons_per_s': m / th, 'photons': m, 'seconds': th,
                     'what': 'CompiledInstrument.process_photons on a host PhotonTable: '
                             'columns and uniforms copied to HBM, FP32 trace, all result '
                             'columns copied back into the table'}
        log('host table path', host_rate)

    # candidate plane tests per launch, counted by the kernels (status
    # counter), and the same step with the rays regrouped between stages
    # (opt-in, identical results)
    tests = {}
    for prec, out in ((32, out32), (64, out64)):
        st1 = torch.zeros(8, dtype=torch.int64, device=dev)
        ci.launch(out, uni, precision=prec, src=src, status=st1, count_tests=True)
        tests[prec] = int(st1.cpu()[7])
    regrouped = {}
    for prec in (32, 64):
        out_r = DevicePhotons(src.dir, src.pol, src.energy, src.prob, device=dev)

        def step_rg(o, p):
            ci.launch(o, uni, precision=p, src=src, status=status, regroup=True)
        wr, kr = timed(out_r, prec, step_rg)
        same = all(torch.equal(getattr(out_r, c), getattr(out32 if prec == 32 else out64, c))
                   for c in ('pos', 'dir', 'pol', 'prob'))
        regrouped[prec] = {'ms_per_step': wr / args.steps * 1e3, 'kernel_ms': kr * 1e3,
                           'pairs_per_round': ci.regroup_pairs(prec, regroup=True),
                           'identical_results': bool(same)}
        del out_r
    log('tests per launch', tests, 'with regrouping', regrouped)

    ms32 = wall32 / args.steps
    ms64 = wall64 / args.steps
    pps32 = n / ms32
    pps64 = n / ms64
    gbs = BYTES_PER_PHOTON * n / k32 / 1e9
    traffic = read_traffic(root)
    res = {
        'metric': METRIC,
        'value': pps32,
        'unit': 'photons/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': wall32 / args.steps * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'fp32',
        'data': 'synthetic',
        'config': {'workload': 'cat_rowland_600: 1e7 photons, line spectrum, 600 CAT facets '
                               '+ 16 CCDs (precision config)',
                   'photons': n, 'facets': len(parts['gratings'].elements),
                   'ccds': len(parts['ccds'].elements), 'culling_grid': bool(args.grid)},
        'roofline': {'bound': 'hbm', 'achieved': gbs, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': gbs / HBM_PEAK_GBS,
                     'traffic': traffic},
        'cpu_baseline': cpu,
        'fp32_with_draws': {'photons_per_s': n / (wall_d / args.steps),
                            'ms_per_step': wall_d / args.steps * 1e3,
                            'uniform_slots': ns,
                            'what': 'the FP32 step with its uniform draws generated on the '
                                    'device (torch Philox, float64) inside the timed region'},
        'fp64_photons_per_s': pps64,
        'fp64_ms_per_step': wall64 / args.steps * 1e3,
        'fp32_over_fp64': pps32 / pps64,
        'fp32_error_vs_fp64': err,
        'kernel_ms_fp32': k32 * 1e3,
        'kernel_ms_fp64': k64 * 1e3,
        'grid_tuned': grid_tuned,
        'with_regrouping': regrouped,
        # plane tests the kernels made (status counter), per second of step
        'photon_element_tests': {
            'per_launch_fp32': tests[32], 'per_launch_fp64': tests[64],
            'per_photon_fp32': tests[32] / n, 'per_photon_fp64': tests[64] / n,
            'per_s_fp32': tests[32] / ms32, 'per_s_fp64': tests[64] / ms64,
            'brute_force_tests_per_photon': n_elem,
            'what': 'candidate ray-plane tests of the culling lists (every element where a '
                    'ray is too oblique for them), lens and CCDs included, counted on the GPU'},
        'valu_fp32': valu_from_counters(root, 32, k32, n),
        'valu_fp64': valu_from_counters(root, 64, k64, n),
        'bytes_per_photon': BYTES_PER_PHOTON,
        # counters accumulate over the warm-up and timed launches of both precisions
        'status_per_launch': {'grid_fallback': int(st[3]) / (2 * (args.warmup + args.steps)),
                              'multi_hit': int(st[4]) / (2 * (args.warmup + args.steps))},
    }
    if cpu is not None:
        res['cpu_baseline_multiprocess'] = cpu_mp
        res['cpu_baseline_selector_loop'] = cpu_loop
    if host_rate is not None:
        res['host_table_path'] = host_rate
    del out32, out64, src, uni, u, energy
    torch.cuda.empty_cache()
    if 'adjoint' in sections:
        res['adjoint'] = bench_adjoint(args, dev)
    if 'stack' in sections:
        res['stack'] = bench_stack(args, dev)
    if 'interp' in sections:
        res['interp'] = bench_interp(args, dev)
    if 'stages' in sections:
        res['stages'] = bench_stages(args, dev)
    if 'pipeline' in sections:
        res['pipeline'] = bench_pipeline(args, dev)
    if 'batched' in sections:
        res['batched'] = bench_batched(args, dev)
    print(json.dumps(res), flush=True)


def bench_adjoint(args, dev):
    """Adjoint workload: 200-facet array, 1e6 photons at 0.65 keV."""
    import numpy as np
    import torch
    import marxs_amd as mx
    from marxs_amd import configs, tolerancing
    from marxs_amd.adjoint import AdjointProblem, central_differences
    from marxs_amd.tracer import CompiledInstrument, DevicePhotons
    t0 = time.time()
    inst, parts = configs.cat_rowland_instrument(200)
    tolerancing.wiggle(parts['gratings'], dx=0.1, dy=0.1, dz=0.1, rx=1.4544e-4, ry=1.4544e-4,
                       rz=1.4544e-4, rng=np.random.default_rng(args.seed + 1))
    n = 1_000_000
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed + 2)
    ci = CompiledInstrument(inst, device=dev)
    u = torch.rand((ci.n_uniform_slots + 1, n), dtype=torch.float64, device=dev, generator=g)
    phi = 2 * np.pi * u[-1]
    dirs = torch.zeros((3, n), dtype=torch.float64, device=dev)
    dirs[0] = -1.
    pol = torch.stack([torch.zeros_like(phi), torch.sin(phi), torch.cos(phi)])
    energy = torch.full((n,), 0.65, dtype=torch.float64, device=dev)
    prob = torch.ones(n, dtype=torch.float64, device=dev)
    uni = u[:ci.n_uniform_slots].contiguous()
    full = DevicePhotons(dirs, pol, energy, prob, device=dev)
    ci.trace_device(full, uniforms=uni)
    front = CompiledInstrument(mx.Sequence(elements=[parts['aperture'], parts['lens']]), device=dev)
    ent = DevicePhotons(dirs, pol, energy, prob, device=dev)
    front.trace_device(ent, uniforms=uni[:2].contiguous())
    sel = (full.icols['order'] == -6) & (full.icols['CCD'] >= 0)
    weight = torch.where(sel, full.prob, torch.zeros_like(full.prob))
    prob_ = AdjointProblem(parts['gratings'], parts['ccds'], ent.pos, ent.dir, ent.energy, weight,
                           full.icols['facet'], full.icols['order'], full.icols['CCD'], coord=1)
    dropped = prob_.restrict_to_traced(full.pos[1])
    log(f'adjoint problem: {int(sel.sum())} order -6 photons on CCDs ({dropped} met two facets), '
        f'set up in {time.time() - t0:.1f}s')

    def step(prec):
        th = prob_.theta0.clone().requires_grad_(True)
        pe = prob_.period0.clone().requires_grad_(True)
        L = prob_.loss(th, pe, precision=prec)
        L.backward()
        return L, th.grad, pe.grad

    # fused path: pose kernel, forward replay with moments, adjoint with
    # seeds from the moments, pose chain rule; captured once, replayed
    out = {}
    buf = prob_.buffers
    for prec in (32, 64):
        graph = prob_.capture(precision=prec)
        buf['theta'].copy_(prob_.theta0)
        buf['period'].copy_(prob_.period0)
        for _ in range(args.warmup):
            graph.replay()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            graph.replay()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        out[prec] = (dt, buf['grad7'].clone(), float(buf['loss'][0]))
        del graph
    # the torch autograd path (FacetLoss over the pose chain), for comparison
    for _ in range(args.warmup):
        step(32)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        step(32)
    torch.cuda.synchronize()
    dt_autograd = (time.perf_counter() - t) / args.steps
    # kernel time of one backward launch (FP32), events on the launch stream
    geom = prob_.geometry(prob_.theta0, prob_.period0).detach()
    seed = torch.where(prob_.weight != 0, prob_.weight / prob_.weight.sum(), torch.zeros_like(prob_.weight))
    ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
    prob_.backward_kernel(geom, seed, 32)
    ev[0].record()
    for _ in range(args.steps):
        prob_.backward_kernel(geom, seed, 32)
    ev[1].record()
    torch.cuda.synchronize()
    k_bw = ev[0].elapsed_time(ev[1]) * 1e-3 / args.steps
    g32, g64 = out[32][1], out[64][1]
    counts = torch.bincount(prob_.facet[prob_.weight != 0].long(), minlength=prob_.F)
    top = torch.argsort(counts, descending=True)[:3].tolist()
    fd = central_differences(prob_, prob_.theta0, prob_.period0, facets=top)
    names = ['dx', 'dy', 'dz', 'rx', 'ry', 'rz', 'period']
    fd_err = {}
    f32_err = {}
    for k, nm in enumerate(names):
        scale = torch.abs(g64[:, k]).max()
        fd_err[nm] = float((torch.abs(g64[top, k] - fd[top, k]) / scale).max())
        f32_err[nm] = float((torch.abs(g32[:, k] - g64[:, k]) / scale).max())
    res = {'workload': '200 CAT facets, 1e6 photons at 0.65 keV, poses N(0,0.1 mm)/N(0,30 arcsec), '
                       'loss = centroid + RMS width of order -6 (y on the CCDs)',
           'photons': n, 'selected_photons': int((prob_.weight != 0).sum()),
           'adjoints_per_s_fp32': 1. / out[32][0], 'adjoints_per_s_fp64': 1. / out[64][0],
           'ms_per_adjoint_fp32': out[32][0] * 1e3, 'ms_per_adjoint_fp64': out[64][0] * 1e3,
           'path': 'fused HIP pose/replay/adjoint kernels, one hipGraph replay per adjoint',
           'ms_per_adjoint_torch_autograd_fp32': dt_autograd * 1e3,
           'loss_mm': out[64][2],
           'backward_kernel_ms_fp32_uncompacted': k_bw * 1e3,
           'input_photons_per_s_fp32': n / out[32][0],
           'replayed_photons_per_adjoint': int(buf['src']['n']),
           'gradients': '200 x 6 poses + 200 periods',
           'fd_rel_err_fp64_top3_facets': fd_err, 'fd_steps': '1e-4 mm, 1e-6 rad, 1e-10 mm',
           'fp32_vs_fp64_rel_err': f32_err}
    log('adjoint', res)
    return res


def gpu_inputs(ci, n, seed, dev, mono=None):
    """The trace workload's synthetic photons, generated on the GPU: line
    spectrum (or all at energy ``mono`` keV), on-axis directions,
    polarization angle 2 pi u; returns (DevicePhotons, uniforms)."""
    import numpy as np
    import torch
    from marxs_amd import configs
    from marxs_amd.tracer import DevicePhotons
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    u = torch.rand((ci.n_uniform_slots + 3, n), dtype=torch.float64, device=dev, generator=g)
    lines = torch.tensor(configs.LINES, dtype=torch.float64, device=dev)
    cont = u[0] < 0.2
    idx = torch.clamp(((u[0] - 0.2) / 0.2).long(), 0, 3)
    energy = torch.where(cont, 0.3 + 1.7 * u[1], lines[idx])
    if mono is not None:
        energy = torch.full_like(energy, float(mono))
    phi = 2 * np.pi * u[2]
    dirs = torch.zeros((3, n), dtype=torch.float64, device=dev)
    dirs[0] = -1.
    pol = torch.stack([torch.zeros_like(phi), torch.sin(phi), torch.cos(phi)])
    src = DevicePhotons(dirs, pol, energy, torch.ones(n, dtype=torch.float64, device=dev),
                        device=dev)
    return src, u[3:].contiguous()


def bench_stack(args, dev):
    """600 CATL1L2Stack facets + 16 CCDs, 1e7 photons."""
    import numpy as np
    import torch
    from marxs_amd import configs, catgrating
    from marxs_amd.tracer import CompiledInstrument, DevicePhotons, raise_for_status
    from marxs_amd.batch import outputs_like
    inst, parts = configs.cat_rowland_instrument(
        args.facets, facet_class=catgrating.CATL1L2Stack,
        facet_kwargs={'l1_transmission': catgrating.l1transtab_path})
    ci = CompiledInstrument(inst, grid=bool(args.grid), device=dev)
    n = args.photons
    src, uni = gpu_inputs(ci, n, args.seed + 5, dev)
    status = torch.zeros(8, dtype=torch.int64, device=dev)
    outs = {p: outputs_like(src) for p in (32, 64)}
    ms = {}
    for prec in (32, 64):
        for _ in range(args.warmup):
            ci.launch(outs[prec], uni, precision=prec, src=src, status=status)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(args.steps):
            ci.launch(outs[prec], uni, precision=prec, src=src, status=status)
        b.record()
        torch.cuda.synchronize()
        ms[prec] = a.elapsed_time(b) / args.steps
    # the other form of each precision: the split trace (two kernels with a
    # photon queue between them; opt-in) beside the one-kernel default; FP64
    # results must be equal bit for bit
    split_ms, split_eq = {}, None
    if ci.tables.split_stage is not None:
        for prec in (32, 64):
            o = outputs_like(src)
            other = not ci.uses_split(precision=prec)
            for _ in range(args.warmup):
                ci.launch(o, uni, precision=prec, src=src, status=status, split=other)
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(args.steps):
                ci.launch(o, uni, precision=prec, src=src, status=status, split=other)
            b.record()
            torch.cuda.synchronize()
            split_ms[prec] = (other, a.elapsed_time(b) / args.steps)
            if prec == 64:
                split_eq = (all(torch.equal(o.icols[c], outs[64].icols[c]) for c in o.icols)
                            and all(torch.equal(torch.nan_to_num(o.fcols[c], nan=7.25),
                                                torch.nan_to_num(outs[64].fcols[c], nan=7.25))
                                    for c in o.fcols)
                            and torch.equal(o.pos, outs[64].pos) and torch.equal(o.prob, outs[64].prob))
            del o
    raise_for_status(status.cpu().numpy())
    o32, o64 = outs[32], outs[64]
    same = ((o32.icols['CCD'] == o64.icols['CCD']) & (o64.icols['CCD'] >= 0)
            & (o32.icols['facet'] == o64.icols['facet'])
            & (o32.icols['order_L1'] == o64.icols['order_L1']))
    dx = (o32.fcols['det_x'] - o64.fcols['det_x'])[same]
    dy = (o32.fcols['det_y'] - o64.fcols['det_y'])[same]
    rms = float(torch.sqrt(torch.mean(dx * dx + dy * dy)) * 1e3)
    res = {'workload': f'cat_rowland_{args.facets} with CATL1L2Stack facets, {n} photons, '
                       'line spectrum',
           'photons_per_s_fp32': n / ms[32] * 1e3, 'photons_per_s_fp64': n / ms[64] * 1e3,
           'kernel_ms_fp32': ms[32], 'kernel_ms_fp64': ms[64],
           'fp32_rms_um_vs_fp64': rms, 'photons_on_ccds': int((o64.icols['CCD'] >= 0).sum()),
           'through_l1_bars': int(((o64.icols['order_L1'] == 0) & (o64.icols['facet'] >= 0)).sum())}
    if split_ms:
        res['trace_form'] = {f'fp{p}': 'split' if not split_ms[p][0] else 'one kernel' for p in (32, 64)}
        res['other_form'] = {
            f'fp{p}': {'form': 'split' if split_ms[p][0] else 'one kernel', 'kernel_ms': split_ms[p][1]}
            for p in (32, 64)}
        res['other_form']['fp64_split_equals_one_kernel_bitwise'] = bool(split_eq)
        res['other_form']['what'] = ('split: the plain kernel up to the first stack hit, the layer programs '
                                     'over the queued photons (mx_trace_split); one kernel: mx_trace')
    if args.stack_cpu_sample > 0:
        os.environ.setdefault('OMP_NUM_THREADS', '1')
        from oracle import trace as OT
        from oracle.describe import describe
        s = min(args.stack_cpu_sample, n)
        ph = {'dir': src.dir[:, :s].cpu().numpy().T.copy(), 'pol': src.pol[:, :s].cpu().numpy().T.copy(),
              'energy': src.energy[:s].cpu().numpy(), 'prob': src.prob[:s].cpu().numpy()}
        t = time.perf_counter()
        ref = OT.trace(describe(inst), ph, uni[:, :s].cpu().numpy())
        tc = time.perf_counter() - t
        eq = all(np.array_equal(ref[c], o64.icols[c][:s].cpu().numpy().astype(np.int64))
                 for c in ('facet', 'order', 'order_L1'))
        rel = float(np.nanmax(np.abs(ref['L2Diffraction'] - o64.fcols['L2Diffraction'][:s].cpu().numpy())
                              / np.maximum(1., np.abs(ref['L2Diffraction']))))
        res['fp64_sample_indices_equal_oracle'] = bool(eq)
        res['fp64_sample_scatter_max_rel_err'] = rel
        res['cpu_baseline'] = {'value': s / tc, 'unit': 'photons/s', 'cores': 1, 'kind': 'port',
                               'sample': f'first {s} photons, numpy oracle, one thread, {tc:.1f} s'}
    log('stack', res)
    return res


def bench_stages(args, dev):
    """The 600-facet workload cut after each stage: kernel time per launch."""
    import torch
    import marxs_amd as mx
    from marxs_amd import configs
    from marxs_amd.tracer import CompiledInstrument
    from marxs_amd.batch import outputs_like
    inst, parts = configs.cat_rowland_instrument(args.facets)
    cuts = {'aperture': ['aperture'], 'aperture+lens': ['aperture', 'lens'],
            'aperture+lens+facets': ['aperture', 'lens', 'gratings'],
            'aperture+lens+ccds': ['aperture', 'lens', 'ccds'],
            'all': ['aperture', 'lens', 'gratings', 'ccds']}
    out = {}
    for name, keys in cuts.items():
        ci = CompiledInstrument(mx.Sequence(elements=[parts[k] for k in keys]), grid=bool(args.grid),
                                device=dev)
        src, uni = gpu_inputs(ci, args.photons, args.seed, dev)
        for prec in (32, 64):
            o = outputs_like(src)
            for _ in range(args.warmup):
                ci.launch(o, uni, precision=prec, src=src)
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(args.steps):
                ci.launch(o, uni, precision=prec, src=src)
            b.record()
            torch.cuda.synchronize()
            out[f'{name} fp{prec} ms'] = a.elapsed_time(b) / args.steps
            log(f'stages {name} FP{prec}: {out[f"{name} fp{prec} ms"]:.3f} ms')
        del src, uni
    return out


def bench_interp(args, dev):
    """600 CAT facets with a blaze-dependent interpolated efficiency table."""
    import numpy as np
    import torch
    from marxs_amd import configs
    from marxs_amd.tracer import CompiledInstrument, raise_for_status
    from marxs_amd.batch import outputs_like
    inst, parts = configs.cat_rowland_instrument(args.facets,
                                                 efficiency=configs.blaze_interp_efficiency())
    ci = CompiledInstrument(inst, grid=bool(args.grid), device=dev)
    n = args.photons
    src, uni = gpu_inputs(ci, n, args.seed + 9, dev)
    status = torch.zeros(8, dtype=torch.int64, device=dev)
    outs = {p: outputs_like(src) for p in (32, 64)}
    ms = {}
    for prec in (32, 64):
        for _ in range(args.warmup):
            ci.launch(outs[prec], uni, precision=prec, src=src, status=status)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(args.steps):
            ci.launch(outs[prec], uni, precision=prec, src=src, status=status)
        b.record()
        torch.cuda.synchronize()
        ms[prec] = a.elapsed_time(b) / args.steps
    raise_for_status(status.cpu().numpy())
    o32, o64 = outs[32], outs[64]
    same = ((o32.icols['CCD'] == o64.icols['CCD']) & (o64.icols['CCD'] >= 0)
            & (o32.icols['facet'] == o64.icols['facet']))
    dx = (o32.fcols['det_x'] - o64.fcols['det_x'])[same]
    dy = (o32.fcols['det_y'] - o64.fcols['det_y'])[same]
    res = {'workload': f'cat_rowland_{args.facets} with a blaze-dependent InterpolateEfficiencyTable '
                       f'(19 wavelengths x 9 blaze angles x 15 orders), {n} photons, line spectrum',
           'photons_per_s_fp32': n / ms[32] * 1e3, 'photons_per_s_fp64': n / ms[64] * 1e3,
           'kernel_ms_fp32': ms[32], 'kernel_ms_fp64': ms[64],
           'fp32_rms_um_vs_fp64': float(torch.sqrt(torch.mean(dx * dx + dy * dy)) * 1e3),
           'order_fp32_equals_fp64_same_facet': float(
               (o32.icols['order'] == o64.icols['order'])[o32.icols['facet'] == o64.icols['facet']]
               .double().mean())}
    if args.stack_cpu_sample > 0:
        os.environ.setdefault('OMP_NUM_THREADS', '1')
        from oracle import trace as OT
        from oracle.describe import describe
        s = min(args.stack_cpu_sample, n)
        ph = {'dir': src.dir[:, :s].cpu().numpy().T.copy(), 'pol': src.pol[:, :s].cpu().numpy().T.copy(),
              'energy': src.energy[:s].cpu().numpy(), 'prob': src.prob[:s].cpu().numpy()}
        t = time.perf_counter()
        ref = OT.trace(describe(inst), ph, uni[:, :s].cpu().numpy())
        tc = time.perf_counter() - t
        res['fp64_sample_indices_equal_oracle'] = bool(all(
            np.array_equal(ref[c], o64.icols[c][:s].cpu().numpy().astype(np.int64))
            for c in ('facet', 'order', 'CCD')))
        res['cpu_baseline'] = {'value': s / tc, 'unit': 'photons/s', 'cores': 1, 'kind': 'port',
                               'sample': f'first {s} photons, numpy oracle, one thread, {tc:.1f} s'}
    log('interp', res)
    return res


def bench_pipeline(args, dev):
    """Mirror scatter, 600 CAT facets, facet frame, Rowland-circle tube;
    then the per-order figures of merit on the device columns."""
    import numpy as np
    import torch
    import marxs_amd as mx
    from marxs_amd import configs, catgrating, analysis
    from marxs_amd.scatter import RadialMirrorScatter
    from marxs_amd.geometry import Cylinder
    from marxs_amd.tracer import CompiledInstrument, raise_for_status
    from marxs_amd.batch import outputs_like
    inst0, parts = configs.cat_rowland_instrument(args.facets)
    rms = RadialMirrorScatter(inplanescatter=2e-5, perpplanescatter=4e-6,
                              position=[12500., 0, 0], zoom=[1, 700, 700])
    # the quarter of the Rowland circle on the dispersed side (negative
    # orders), from just past the focus
    tube = mx.CircularDetector(geometry=Cylinder.from_rowland(
        parts['rowland'], width=50., rotation=np.pi, kwargs={'phi_lim': [-np.pi / 2, 0.02]}),
        pixsize=0.024)
    inst = mx.Sequence(elements=[parts['aperture'], parts['lens'], rms, parts['gratings'],
                                 catgrating.catsupportbars, tube])
    ci = CompiledInstrument(inst, grid=bool(args.grid), device=dev)
    n = args.photons
    src, uni = gpu_inputs(ci, n, args.seed + 7, dev, mono=0.65)
    status = torch.zeros(8, dtype=torch.int64, device=dev)
    outs = {p: outputs_like(src) for p in (32, 64)}
    ms = {}
    for prec in (32, 64):
        for _ in range(args.warmup):
            ci.launch(outs[prec], uni, precision=prec, src=src, status=status)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(args.steps):
            ci.launch(outs[prec], uni, precision=prec, src=src, status=status)
        b.record()
        torch.cuda.synchronize()
        ms[prec] = a.elapsed_time(b) / args.steps
    raise_for_status(status.cpu().numpy())
    o64 = outs[64]
    orders = np.arange(-12, 3)
    area = np.pi * (600. ** 2 - 300. ** 2)
    cap = analysis.CaptureResAeff(A_geom=area, order_col='order', orders=orders,
                                  dispersion_coord='det_phi')
    out = cap(o64)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        out = cap(o64)
    t_an = (time.perf_counter() - t) / args.steps
    host = {'order': o64.icols['order'].cpu().numpy(), 'det_phi': o64.fcols['det_phi'].cpu().numpy(),
            'prob': o64.prob.cpu().numpy()}
    from oracle import analysis as OA
    t = time.perf_counter()
    ref = OA.capture_res_aeff(host['order'], host['det_phi'], host['prob'], orders, A_geom=area)
    t_cpu = time.perf_counter() - t
    fin = np.isfinite(ref['R']) & (ref['R'] > 0)
    res = {'workload': f'aperture, lens, RadialMirrorScatter, cat_rowland_{args.facets} facets, '
                       f'catsupportbars, Rowland-circle CircularDetector; {n} photons at 0.65 keV',
           'photons_per_s_fp32': n / ms[32] * 1e3, 'photons_per_s_fp64': n / ms[64] * 1e3,
           'kernel_ms_fp32': ms[32], 'kernel_ms_fp64': ms[64],
           'photons_on_tube': int(torch.isfinite(o64.fcols['det_phi']).sum()),
           'capture_res_aeff_ms_gpu': t_an * 1e3,
           'capture_res_aeff_ms_cpu_oracle': t_cpu * 1e3,
           'capture_photons_per_s_gpu': n / t_an,
           'R_orders_finite': int(fin.sum()),
           'R_max_rel_diff_vs_oracle': float(np.max(np.abs(out['R'][fin] - ref['R'][fin])
                                                    / np.abs(ref['R'][fin]))) if fin.any() else None,
           'Aeff_max_rel_diff_vs_oracle': float(np.max(np.abs(out['Aeff'] - ref['Aeff'])
                                                       / np.maximum(ref['Aeff'], 1e-300))),
           'Rgrat': float(out['Rgrat']), 'Aeffgrat_mm2': float(out['Aeffgrat'])}
    log('pipeline', res)
    return res


def bench_batched(args, dev):
    """64 design variants traced in one launch vs one launch per variant."""
    import numpy as np
    import torch
    from marxs_amd import configs
    from marxs_amd.batch import BatchedTrace, outputs_like
    from marxs_amd.autotune import tune_batched
    from marxs_amd.tracer import CompiledInstrument, DevicePhotons
    t0 = time.time()
    variants = configs.design_variants(args.variants, np.random.default_rng(args.seed + 3))
    cis = [CompiledInstrument(v[0], device=dev) for v in variants]
    log(f'{len(cis)} variants ({sum(v[2]["n_facets"] for v in variants)} facets) built and '
        f'compiled in {time.time() - t0:.1f}s')
    n = args.variant_photons
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed + 4)
    inputs, unis = [], []
    for ci in cis:
        u = torch.rand((ci.n_uniform_slots + 2, n), dtype=torch.float64, device=dev, generator=g)
        phi = 2 * np.pi * u[-1]
        dirs = torch.zeros((3, n), dtype=torch.float64, device=dev)
        dirs[0] = -1.
        pol = torch.stack([torch.zeros_like(phi), torch.sin(phi), torch.cos(phi)])
        energy = 0.3 + 1.7 * u[-2]
        inputs.append(DevicePhotons(dirs, pol, energy, torch.ones(n, dtype=torch.float64, device=dev),
                                    device=dev))
        unis.append(u[:ci.n_uniform_slots].contiguous())
    total = n * len(cis)
    res = {'workload': f'{len(cis)} design variants (R ~ U(5000,7000) mm, blaze ~ U(1,2) deg, '
                       f'50-800 facets, 16 CCDs each), {n} photons each, 0.3-2.0 keV',
           'variants': len(cis), 'photons': total,
           'facets': [v[2]['n_facets'] for v in variants]}
    for prec in (32, 64):
        bt = BatchedTrace(cis, inputs, unis, precision=prec)
        best, table = tune_batched(bt)
        bt.run()
        bt.capture()
        for _ in range(args.warmup):
            bt.replay()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            bt.replay()
        torch.cuda.synchronize()
        tb = (time.perf_counter() - t) / args.steps
        bt.run(use_graph=True)
        # one launch per variant, same tables and photons
        seq_out = [outputs_like(src) for src in inputs]

        def seq():
            for ci, src, u, o in zip(cis, inputs, unis, seq_out):
                ci.launch(o, u, prec, src=src)
        for _ in range(args.warmup):
            seq()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            seq()
        torch.cuda.synchronize()
        ts = (time.perf_counter() - t) / args.steps
        r = {'photons_per_s_batched': total / tb, 'ms_per_batch': tb * 1e3,
             'photons_per_s_sequential': total / ts, 'ms_sequential': ts * 1e3,
             'speedup_vs_sequential': ts / tb,
             'plan': {k: (v if k != 'tile_photons' else v) for k, v in best.items()},
             'autotune_ms': {f"{c['tile_photons']}/{'scaled' if c['scale_by_cost'] else 'flat'}/"
                             f"{c['order'] if isinstance(c['order'], str) else 'perm'}/"
                             f"wg{c['wg_per_cu']}": ms for c, ms in table}}
        if prec == 64:
            ident = True
            for o, q in zip(bt.outputs, seq_out):
                for a, b in [(o.pos, q.pos), (o.dir, q.dir), (o.pol, q.pol), (o.prob, q.prob)] + \
                        [(o.fcols[k], q.fcols[k]) for k in o.fcols] + \
                        [(o.icols[k], q.icols[k]) for k in o.icols]:
                    if not torch.equal(torch.nan_to_num(a, nan=12345.), torch.nan_to_num(b, nan=12345.)):
                        ident = False
            r['fp64_batched_equals_sequential_bitwise'] = ident
        res[f'fp{prec}'] = r
        log(f'batched FP{prec}', r)
        del bt, seq_out
    return res


if __name__ == '__main__':
    main()
