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
oracle's rate on a bounded sample are reported beside it.

    python bench.py [--steps K] [--warmup W] [--photons N] [--cpu-sample S]
"""
import argparse
import json
import os
import re
import sys
import time

HBM_PEAK_GBS = 8000.0
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
    ap.add_argument('--cpu-sample', type=int, default=200_000)
    ap.add_argument('--seed', type=int, default=20261202)
    ap.add_argument('--grid', type=int, default=1)
    ap.add_argument('--only', choices=['fp32', 'fp64'], default=None,
                    help='profiling: time one precision only, skip comparisons')
    ap.add_argument('--copy-calibration-mib', type=int, default=0,
                    help='profiling: run one device copy of this size first (counter calibration)')
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(root):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by a copy-kernel calibration), committed in
    profiles/trace_c3_fp32_pmc.txt as 'traffic_bytes_per_launch: X'."""
    path = os.path.join(root, 'profiles', 'trace_c3_fp32_pmc.txt')
    if not os.path.exists(path):
        return None
    m = re.search(r'traffic_bytes_per_launch:\s*([0-9.eE+]+)', open(path).read())
    return float(m.group(1)) if m else None


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

    def step(out, prec):
        ci.launch(out, uni, precision=prec, src=src, status=status)

    def timed(out, prec):
        for _ in range(args.warmup):
            step(out, prec)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
        t = time.perf_counter()
        for a, b in ev:
            a.record()
            step(out, prec)
            b.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        kern = [a.elapsed_time(b) * 1e-3 for a, b in ev]
        return wall, float(np.mean(kern))

    if args.copy_calibration_mib:
        a = torch.empty(args.copy_calibration_mib * 2 ** 20 // 8, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log(f'copy calibration: {a.numel() * 8} bytes read, {a.numel() * 8} bytes written')
        del a, b
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

    pps32 = n / k32
    pps64 = n / k64
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
        'fp64_photons_per_s': pps64,
        'fp64_ms_per_step': wall64 / args.steps * 1e3,
        'fp32_over_fp64': pps32 / pps64,
        'fp32_error_vs_fp64': err,
        'photon_element_tests_per_s_fp32': pps32 * n_elem,
        'photon_element_tests_per_s_fp64': pps64 * n_elem,
        'kernel_ms_fp32': k32 * 1e3,
        'kernel_ms_fp64': k64 * 1e3,
        'bytes_per_photon': BYTES_PER_PHOTON,
        'status': {'grid_fallback': int(st[3]), 'multi_hit': int(st[4])},
    }
    print(json.dumps(res), flush=True)


if __name__ == '__main__':
    main()
