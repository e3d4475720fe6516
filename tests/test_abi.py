"""The C-ABI library: loads, exports every symbol of include/marxs_hip.h and
matches the ctypes struct layout (CPU only; no compute calls)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'marxs_hip.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(mx_[a-z0-9_]+)\s*\(', text)))


@pytest.fixture(scope='module')
def lib():
    from marxs_amd import _lib
    return _lib


def test_library_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert 'mx_trace' in names
    for n in names:
        assert hasattr(lib.lib, n), n
    out = subprocess.run(['nm', '-D', '--defined-only', lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r' T (mx_\w+)', out))
    assert set(names) <= exported


def test_abi_version(lib):
    assert lib.lib.mx_abi_version() == lib.MX_ABI_VERSION


def test_struct_layout_matches_header(lib, tmp_path):
    """sizeof/offsetof from the C compiler against the ctypes mirror."""
    fields = {'MxTables': [f[0] for f in lib.MxTables._fields_],
              'MxPhotons': [f[0] for f in lib.MxPhotons._fields_],
              'MxAdjointArgs': [f[0] for f in lib.MxAdjointArgs._fields_],
              'MxBatch': [f[0] for f in lib.MxBatch._fields_],
              'MxPoseArgs': [f[0] for f in lib.MxPoseArgs._fields_],
              'MxOrderStats': [f[0] for f in lib.MxOrderStats._fields_]}
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "marxs_hip.h"',
           'int main(void) {']
    for st, fs in fields.items():
        src.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for f in fs:
            src.append(f'printf("{st}.{f} %zu\\n", offsetof({st}, {f}));')
    src.append('return 0; }')
    c = tmp_path / 'layout.c'
    c.write_text('\n'.join(src))
    exe = tmp_path / 'layout'
    subprocess.run(['gcc', '-I', os.path.join(ROOT, 'include'), str(c), '-o', str(exe)],
                   check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True,
                                                       text=True, check=True).stdout.splitlines())
    for st, fs in fields.items():
        cls = getattr(lib, st)
        assert int(got[st]) == ctypes.sizeof(cls)
        for f in fs:
            assert int(got[f'{st}.{f}']) == getattr(cls, f).offset, (st, f)


def test_argument_errors_need_no_gpu(lib):
    """Invalid arguments are rejected on the host before any launch."""
    t = lib.MxTables()
    p = lib.MxPhotons()
    p.n = -1
    rc = lib.lib.mx_trace(ctypes.byref(t), ctypes.byref(p), 64, 0, None, None)
    assert rc != 0
    assert b'negative' in lib.lib.mx_last_error()
    p.n = 0
    rc = lib.lib.mx_trace(ctypes.byref(t), ctypes.byref(p), 16, 0, None, None)
    assert rc != 0 and b'precision' in lib.lib.mx_last_error()


def test_split_workspace_layout_and_errors_need_no_gpu(lib):
    """mx_trace_split_workspace: 256 region counters of 128 bytes, then
    photon index, hit point (3), direction (3) and polarization (3) at the hit
    (float32 on the FP32 path), probability, two local coordinates and the
    element index, each array 256-byte aligned, for 256 regions of whole
    blocks of the first kernel (1024 photons FP32, 768 FP64)."""
    def a(b):
        return (b + 255) // 256 * 256
    for n in (0, 1, 1000, 10_000_019):
        for prec, ev, block in ((64, 8, 768), (32, 4, 1024)):
            nb = -(-n // block)
            cap = -(-nb // 256) * block * 256
            assert cap >= n
            want = 128 * 256 + a(8 * cap) * 4 + a(ev * cap) * 6 + a(8 * cap) * 3 + a(4 * cap)
            assert lib.lib.mx_trace_split_workspace(n, prec) == want
    assert lib.lib.mx_trace_split_workspace(-1, 64) < 0
    assert b'negative' in lib.lib.mx_last_error()
    assert lib.lib.mx_trace_split_workspace(10, 16) < 0
    t = lib.MxTables()
    p = lib.MxPhotons()
    assert lib.lib.mx_trace_split(ctypes.byref(t), ctypes.byref(p), 16, 0, 0, None, 0, None, None) != 0
    assert b'precision' in lib.lib.mx_last_error()
    assert lib.lib.mx_trace_split(ctypes.byref(t), ctypes.byref(p), 32, 0, 0, None, 0, None, None) == 0


def test_batched_argument_errors_need_no_gpu(lib):
    b = lib.MxBatch()
    assert lib.lib.mx_trace_batched(None, 64, 0, None) != 0
    assert b'NULL' in lib.lib.mx_last_error()
    b.n_variants = -1
    assert lib.lib.mx_trace_batched(ctypes.byref(b), 64, 0, None) != 0
    assert b'negative' in lib.lib.mx_last_error()
    b.n_variants = 0
    assert lib.lib.mx_trace_batched(ctypes.byref(b), 16, 0, None) != 0
    assert b'precision' in lib.lib.mx_last_error()
    assert lib.lib.mx_trace_batched(ctypes.byref(b), 32, 0, None) == 0
    b.n_variants = 2
    assert lib.lib.mx_trace_batched(ctypes.byref(b), 32, 0, None) != 0
    assert b'NULL' in lib.lib.mx_last_error()


def test_order_stats_argument_errors_need_no_gpu(lib):
    a = lib.MxOrderStats()
    assert lib.lib.mx_order_stats(None, None) != 0
    a.n, a.n_orders = 10, 0
    assert lib.lib.mx_order_stats(ctypes.byref(a), None) != 0
    assert b'n_orders' in lib.lib.mx_last_error()
    a.n_orders = 3
    assert lib.lib.mx_order_stats(ctypes.byref(a), None) != 0
    assert b'NULL' in lib.lib.mx_last_error()
    a.n = -1
    assert lib.lib.mx_order_stats(ctypes.byref(a), None) != 0
    assert lib.lib.mx_order_stats_workspace(-1, 3, 0) == -1
    assert lib.lib.mx_order_stats_workspace(10, 65, 0) == -1


def test_adjoint_argument_errors_need_no_gpu(lib):
    """mx_adjoint / mx_facet_pose reject bad flags and modes on the host
    (the pointers below are never dereferenced: validation fails first)."""
    a = lib.MxAdjointArgs()
    a.n, a.n_facet, a.n_ccd, a.precision = 10, 2, 1, 64
    fake = 4096
    for k in range(3):
        a.pos[k] = a.dir[k] = fake
    a.energy = a.facet = a.order = a.ccd = a.geom = a.ccd_t = fake
    a.flags = 2
    assert lib.lib.mx_adjoint(ctypes.byref(a), None) != 0
    assert b'flags' in lib.lib.mx_last_error()
    a.flags = lib.MX_AJ_SUMS
    assert lib.lib.mx_adjoint(ctypes.byref(a), None) != 0       # no grad, no moments
    assert b'MX_AJ_SUMS' in lib.lib.mx_last_error()
    a.grad = fake
    assert lib.lib.mx_adjoint(ctypes.byref(a), None) != 0       # no moments
    assert b'MX_AJ_SUMS' in lib.lib.mx_last_error()
    p = lib.MxPoseArgs()
    p.n_facet, p.mode = 2, 3
    assert lib.lib.mx_facet_pose(ctypes.byref(p), None) != 0
    assert b'mode' in lib.lib.mx_last_error()
    p.mode = 2
    p.theta = p.gu = p.t0 = p.rpost = p.zoom = p.grad13 = p.grad7 = fake
    assert lib.lib.mx_facet_pose(ctypes.byref(p), None) != 0
    assert b'moments' in lib.lib.mx_last_error()
