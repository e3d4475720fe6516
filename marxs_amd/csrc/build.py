"""Build libmarxs_hip.so for gfx950 with hipcc (run on the host; no GPU needed).

    python -m marxs_amd.csrc.build [--verbose]

The library is written next to the package (``marxs_amd/libmarxs_hip.so``)
so that it travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
SOURCES = ['marxs_trace.hip', 'marxs_adjoint.hip', 'marxs_analysis.hip']
TARGET = os.path.join(PKG, 'libmarxs_hip.so')

FLAGS = [
    '-O3', '-std=c++17', '-fPIC', '-shared',
    '--offload-arch=gfx950',
    # parity path: no FMA contraction unless a function opts in
    '-ffp-contract=off',
    # packed FP32 VALU ops stay off
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-I', os.path.join(ROOT, 'include'),
]


def hipcc():
    rocm = os.environ.get('ROCM_PATH', '/opt/rocm')
    exe = os.path.join(rocm, 'bin', 'hipcc')
    return exe if os.path.exists(exe) else 'hipcc'


def build(verbose=False, extra=()):
    srcs = [os.path.join(HERE, s) for s in SOURCES if os.path.exists(os.path.join(HERE, s))]
    tmp = TARGET + '.tmp'
    cmd = [hipcc()] + FLAGS + list(extra) + srcs + ['-o', tmp]
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError('hipcc failed building libmarxs_hip.so')
    if verbose and res.stderr.strip():
        sys.stderr.write(res.stderr)
    os.replace(tmp, TARGET)
    return TARGET


if __name__ == '__main__':
    print(build(verbose='--verbose' in sys.argv))
