"""ctypes binding of libmarxs_hip.so (declarations in include/marxs_hip.h).

The library is loaded after torch so that it shares torch's HIP runtime
(both carry the soname libamdhip64.so.7).  There is no fallback: if the
library is missing the import fails with instructions to build it.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime first)

__all__ = ['lib', 'MxTables', 'MxPhotons', 'MxAdjointArgs', 'MxBatch', 'MxPoseArgs', 'MxOrderStats',
           'MX_MAX_FCOLS', 'MX_MAX_ICOLS', 'MX_NSTATUS', 'MX_MAX_ORDERS', 'MX_OS_FILTER', 'LIB_PATH',
           'check']

MX_MAX_FCOLS = 16
MX_MAX_ICOLS = 8
MX_NSTATUS = 8
MX_ABI_VERSION = 21

LIB_PATH = os.environ.get('MARXS_AMD_LIB') or os.path.join(
    os.path.dirname(os.path.abspath(__file__)), 'libmarxs_hip.so')

P = ctypes.c_void_p


class MxTables(ctypes.Structure):
    _fields_ = [('n_stages', ctypes.c_int32), ('n_elems', ctypes.c_int32),
                ('lds_bytes32', ctypes.c_int32), ('lds_bytes64', ctypes.c_int32),
                ('stage_i', P), ('stage_d', P), ('stage_f', P), ('elem_t', P), ('elem_p', P),
                ('elem_i', P), ('elem_t32', P), ('elem_p32', P), ('sel_i', P),
                ('sel_d', P), ('sel_o', P), ('cell_start', P), ('cell_items', P),
                ('succ_start', P), ('succ_items', P), ('lds_image32', P), ('lds_image64', P),
                ('first_draw_slot', ctypes.c_int32), ('ext_stages', ctypes.c_int32),
                ('regroup_stages', ctypes.c_int32), ('blaze_stages', ctypes.c_int32)]


class MxPhotons(ctypes.Structure):
    _fields_ = [('n', ctypes.c_int64), ('read_pos', ctypes.c_int32),
                ('n_uniform_slots', ctypes.c_int32), ('uniform_stride', ctypes.c_int64),
                ('pos', P * 3), ('dir', P * 3), ('pol', P * 3),
                ('energy', P), ('prob', P), ('uniforms', P),
                ('fcol', P * MX_MAX_FCOLS), ('icol', P * MX_MAX_ICOLS),
                ('pos_in', P * 3), ('dir_in', P * 3), ('pol_in', P * 3), ('prob_in', P)]


class MxAdjointArgs(ctypes.Structure):
    _fields_ = [('n', ctypes.c_int64), ('pos', P * 3), ('dir', P * 3), ('energy', P),
                ('weight', P), ('facet', P), ('order', P), ('ccd', P), ('geom', P),
                ('ccd_t', P), ('seed', P), ('y_out', P), ('grad', P), ('moments', P),
                ('seed_moments', P), ('y_shift', ctypes.c_double),
                ('n_facet', ctypes.c_int32), ('n_ccd', ctypes.c_int32),
                ('coord', ctypes.c_int32), ('precision', ctypes.c_int32),
                ('lds_accumulate', ctypes.c_int32), ('flags', ctypes.c_int32)]


class MxBatch(ctypes.Structure):
    _fields_ = [('n_variants', ctypes.c_int32), ('grid', ctypes.c_int32),
                ('n_tiles', ctypes.c_int64), ('tables', P), ('photons', P),
                ('tables_dev', P), ('photons_dev', P), ('tile_first', P),
                ('tile_photons', P), ('queue', P), ('status', P)]


class MxPoseArgs(ctypes.Structure):
    _fields_ = [('n_facet', ctypes.c_int32), ('mode', ctypes.c_int32), ('theta', P),
                ('period', P), ('gu', P), ('t0', P), ('rpost', P), ('zoom', P), ('extra', P),
                ('geom', P), ('grad13', P), ('grad7', P), ('moments', P), ('loss', P),
                ('y_shift', ctypes.c_double)]


class MxOrderStats(ctypes.Structure):
    _fields_ = [('n', ctypes.c_int64), ('order', P), ('value', P), ('prob', P), ('orders', P),
                ('n_orders', ctypes.c_int32), ('flags', ctypes.c_int32), ('sigma', ctypes.c_double),
                ('maxiters', ctypes.c_int32), ('reserved', ctypes.c_int32), ('workspace', P),
                ('workspace_bytes', ctypes.c_int64), ('prob_sum', P), ('count', P), ('mean', P),
                ('median', P), ('std', P), ('n_kept', P)]


MX_AJ_GEOM = 16
MX_AJ_NGRAD = 13
MX_AJ_SUMS = 1
MX_MAX_ORDERS = 64
MX_OS_FILTER = 1
MX_OS_STATIC_BYTES = 64 << 20


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f'{LIB_PATH} not found: build it with `python -m marxs_amd.csrc.build` '
            '(or __graft_entry__.build()).  The tracer has no CPU implementation.')
    lib = ctypes.CDLL(LIB_PATH)
    lib.mx_abi_version.restype = ctypes.c_int
    lib.mx_last_error.restype = ctypes.c_char_p
    lib.mx_trace.restype = ctypes.c_int
    lib.mx_trace.argtypes = [ctypes.POINTER(MxTables), ctypes.POINTER(MxPhotons),
                             ctypes.c_int32, ctypes.c_int32, P, P]
    lib.mx_trace_regroup_pairs.restype = ctypes.c_int
    lib.mx_trace_regroup_pairs.argtypes = [ctypes.POINTER(MxTables), ctypes.c_int32, ctypes.c_int32]
    lib.mx_trace_split_workspace.restype = ctypes.c_int64
    lib.mx_trace_split_workspace.argtypes = [ctypes.c_int64, ctypes.c_int32]
    lib.mx_trace_split.restype = ctypes.c_int
    lib.mx_trace_split.argtypes = [ctypes.POINTER(MxTables), ctypes.POINTER(MxPhotons), ctypes.c_int32,
                                   ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64,
                                   ctypes.c_void_p, ctypes.c_void_p]
    lib.mx_fill_i32.restype = ctypes.c_int
    lib.mx_fill_i32.argtypes = [P, ctypes.c_int64, ctypes.c_int32, P]
    lib.mx_trace_batched.restype = ctypes.c_int
    lib.mx_trace_batched.argtypes = [ctypes.POINTER(MxBatch), ctypes.c_int32, ctypes.c_int32, P]
    lib.mx_adjoint.restype = ctypes.c_int
    lib.mx_adjoint.argtypes = [ctypes.POINTER(MxAdjointArgs), P]
    lib.mx_facet_pose.restype = ctypes.c_int
    lib.mx_facet_pose.argtypes = [ctypes.POINTER(MxPoseArgs), P]
    lib.mx_order_stats_workspace.restype = ctypes.c_int64
    lib.mx_order_stats_workspace.argtypes = [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32]
    lib.mx_order_stats.restype = ctypes.c_int
    lib.mx_order_stats.argtypes = [ctypes.POINTER(MxOrderStats), P]
    lib.mx_order_stats_static.restype = ctypes.c_int
    lib.mx_order_stats_static.argtypes = [ctypes.POINTER(MxOrderStats), P]
    v = lib.mx_abi_version()
    if v != MX_ABI_VERSION:
        raise ImportError(f'libmarxs_hip.so ABI {v}, expected {MX_ABI_VERSION}: rebuild it')
    return lib


lib = _load()


def check(rc):
    if rc != 0:
        raise RuntimeError('libmarxs_hip: ' + lib.mx_last_error().decode(errors='replace'))
