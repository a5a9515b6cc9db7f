"""ctypes binding of libwql.so (the C ABI of include/wql.h).

The HIP library is the only compute path: importing this module without a
built libwql.so raises, there is no CPU fallback.
"""
import ctypes as C
import os

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, 'libwql.so')

dp = C.POINTER(C.c_double)

WQL_FP64, WQL_FP32 = 0, 1

# error codes of include/wql.h
WQL_E_ARG, WQL_E_NOMEM, WQL_E_HIP, WQL_E_NONFINITE, WQL_E_SOLVER, WQL_E_STATE = -1, -2, -3, -4, -5, -6

EXPORTS = ['wql_create', 'wql_destroy', 'wql_last_error', 'wql_abi_version', 'wql_device_bytes',
           'wql_estimate_bytes', 'wql_init_state', 'wql_set_fields', 'wql_get_fields', 'wql_get_rates',
           'wql_set_pml', 'wql_get_pml', 'wql_get_pml_rates', 'wql_get_metrics', 'wql_get_frame',
           'wql_set_fault', 'wql_get_fault', 'wql_get_fault_rates', 'wql_eval_rates', 'wql_step',
           'wql_sync', 'wql_time', 'wql_get_fault_peak', 'wql_fault_stats', 'wql_server_create', 'wql_server_destroy',
           'wql_server_submit', 'wql_server_poll', 'wql_server_result', 'wql_server_status', 'wql_server_set_policy', 'wql_set_tile', 'wql_get_tile', 'wql_set_tile_order', 'wql_get_tile_order', 'wql_tile_variants', 'wql_use_graph', 'wql_profile',
           'wql_volume_kernel_ms', 'wql_get_fault_faces', 'wql_get_plastic',
           'wql_set_anelastic', 'wql_get_anelastic',
           'wql_get_fields_noalloc', 'wql_problem_keys', 'wql_grid_block', 'wql_problem_material',
           'wql_grid_faces', 'wql_set_sources_smooth',
           'wql_stage', 'wql_server_pool_stats', 'wql_server_set_shape', 'wql_find_stations',
           'wql_set_stations', 'wql_get_stations', 'wql_locate_sources', 'wql_set_sources']


class Source(C.Structure):
    _fields_ = [('type', C.c_int), ('duration', C.c_double), ('t_init', C.c_double), ('m', C.c_double * 6),
                ('x', C.c_double * 3), ('node', C.c_int * 3), ('node_x', C.c_double * 3)]


class FaultParams(C.Structure):
    _fields_ = [('law', C.c_int), ('stress_mode', C.c_int), ('T0', C.c_double * 3), ('s0', C.c_double * 6),
                ('npatch', C.c_int), ('patch', (C.c_double * 5) * 4), ('nuc', C.c_int), ('nucp', C.c_double * 5),
                ('fs', C.c_double), ('fd', C.c_double), ('Dc', C.c_double),
                ('bar_wy', C.c_double), ('bar_wz', C.c_double), ('bar_big', C.c_double),
                ('rs_a', C.c_double), ('rs_da', C.c_double), ('rs_w', C.c_double), ('rs_wr', C.c_double),
                ('rs_b', C.c_double), ('rs_L', C.c_double), ('rs_V0', C.c_double), ('rs_f0', C.c_double),
                ('rs_Vin', C.c_double),
                ('rs_law', C.c_int), ('rs_fw', C.c_double), ('rs_Vw', C.c_double), ('rs_dVw', C.c_double),
                ('rs_init', C.c_int), ('rs_W0', C.c_double), ('sw_law', C.c_int), ('coh', C.c_double * 3),
                ('frc', C.c_double * 5), ('rs_solver', C.c_int), ('key', C.c_int)]


# wql_fault_params.key: slip-weakening problems with per-node fault fields
FAULT_KEYS = {'TPV10': 1, 'TPV29': 2, 'TPV30': 2, 'TPV31': 3, 'TPV32': 4, 'TPV33': 5, 'TPV34': 6,
              'SCITS2016': 7, 'TPV36': 8, 'TPV37': 9, 'wasatch100': 10}


class GridDesc(C.Structure):
    _fields_ = [('nq', C.c_int), ('nr', C.c_int), ('ns', C.c_int), ('a', C.c_double * 3), ('b', C.c_double * 3),
                ('mesh', C.c_int), ('profile', C.c_int), ('lc', C.c_double), ('rc', C.c_double), ('fault_x', dp),
                ('fault_side', C.c_int), ('top_y', dp)]


class BlockDesc(C.Structure):
    _fields_ = [('nq', C.c_int), ('nr', C.c_int), ('ns', C.c_int), ('bc', C.c_int * 6), ('pml', C.c_int * 6),
                ('npml', C.c_int), ('cartesian', C.c_int), ('X', dp), ('M', dp), ('rho_s_p', C.c_double * 3)]


class ProblemDesc(C.Structure):
    _fields_ = [('precision', C.c_int), ('dt', C.c_double), ('t0', C.c_double), ('plastic', C.c_int),
                ('pl_c', C.c_double), ('pl_nu', C.c_double), ('pl_Tv', C.c_double),
                ('pl_mu', C.c_double), ('pl_beta', C.c_double), ('pl_eta', C.c_double),
                ('s0_kind', C.c_int), ('fault', FaultParams), ('fd_type', C.c_int), ('fd_order', C.c_int),
                ('anelastic', C.c_int), ('an_c', C.c_double), ('an_weight_exp', C.c_double),
                ('an_fref', C.c_double)]


FD_TYPES = {'traditional': 0, 'upwind': 1, 'upwind_drp': 2}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f'{LIB_PATH} is not built; run __graft_entry__.build()')
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sig = {
            'wql_create': ([C.POINTER(ProblemDesc), C.POINTER(BlockDesc), C.POINTER(vp)], C.c_int),
            'wql_destroy': ([vp], C.c_int),
            'wql_last_error': ([], C.c_char_p),
            'wql_abi_version': ([], C.c_int),
            'wql_device_bytes': ([vp], C.c_size_t),
            'wql_estimate_bytes': ([C.POINTER(ProblemDesc), C.POINTER(BlockDesc)], C.c_size_t),
            'wql_init_state': ([vp], C.c_int),
            'wql_set_fields': ([vp, C.c_int, dp], C.c_int),
            'wql_get_fields': ([vp, C.c_int, dp], C.c_int),
            'wql_get_rates': ([vp, C.c_int, dp], C.c_int),
            'wql_set_pml': ([vp, C.c_int, C.c_int, dp], C.c_int),
            'wql_get_pml': ([vp, C.c_int, C.c_int, dp], C.c_int),
            'wql_get_pml_rates': ([vp, C.c_int, C.c_int, dp], C.c_int),
            'wql_get_metrics': ([vp, C.c_int, dp], C.c_int),
            'wql_get_frame': ([vp, C.c_int, C.c_int, dp], C.c_int),
            'wql_set_fault': ([vp, C.c_int, dp, dp], C.c_int),
            'wql_get_fault': ([vp, C.c_int, dp, dp, dp, dp, dp, dp], C.c_int),
            'wql_get_fault_rates': ([vp, C.c_int, dp, dp], C.c_int),
            'wql_get_plastic': ([vp, C.c_int, dp], C.c_int),
            'wql_set_anelastic': ([vp, C.c_int, dp], C.c_int),
            'wql_get_anelastic': ([vp, C.c_int, C.c_int, dp], C.c_int),
            'wql_get_fields_noalloc': ([vp, C.c_int, dp], C.c_int),
            'wql_problem_material': ([C.c_char_p, C.c_int, C.c_int, C.c_int, dp, dp], C.c_int),
            'wql_grid_faces': ([C.c_int, C.c_int, C.c_int, dp, dp], C.c_int),
            'wql_problem_keys': ([C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, dp, C.POINTER(ProblemDesc)],
                                 C.c_int),
            'wql_grid_block': ([C.POINTER(GridDesc), dp], C.c_int),
            'wql_eval_rates': ([vp, C.c_double, C.c_int, C.c_int], C.c_int),
            'wql_step': ([vp, C.c_int], C.c_int),
            'wql_stage': ([vp, C.c_int], C.c_int),
            'wql_sync': ([vp], C.c_int),
            'wql_time': ([vp], C.c_double),
            'wql_get_fault_peak': ([vp, C.c_int, dp], C.c_int),
            'wql_get_fault_faces': ([vp, dp, dp], C.c_int),
            'wql_server_create': ([C.c_int, C.c_size_t, C.c_int, C.POINTER(vp)], C.c_int),
            'wql_server_destroy': ([vp], C.c_int),
            'wql_server_submit': ([vp, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)], C.c_int),
            'wql_server_poll': ([vp, C.POINTER(C.c_int), C.c_int], C.c_int),
            'wql_server_result': ([vp, C.c_int, dp, dp, dp, dp], C.c_int),
            'wql_server_status': ([vp, C.POINTER(C.c_long)], C.c_int),
            'wql_server_pool_stats': ([vp, C.POINTER(C.c_long)], C.c_int),
            'wql_server_set_policy': ([vp, C.c_int], C.c_int),
            'wql_find_stations': ([dp, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.POINTER(C.c_int)], C.c_int),
            'wql_set_stations': ([vp, C.c_int, C.c_int, C.POINTER(C.c_int)], C.c_int),
            'wql_locate_sources': ([dp, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.POINTER(Source)], C.c_int),
            'wql_set_sources': ([vp, C.c_int, C.c_int, C.POINTER(Source), C.c_int, C.c_int], C.c_int),
            'wql_set_sources_smooth': ([vp, C.c_int, C.c_int, C.POINTER(Source), dp], C.c_int),
            'wql_get_stations': ([vp, C.c_int, dp], C.c_int),
            'wql_server_set_shape': ([vp, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
                                     C.c_int),
            'wql_fault_stats': ([vp, C.c_int, dp], C.c_int),
            'wql_set_tile': ([vp, C.c_int, C.c_int], C.c_int),
            'wql_get_tile': ([vp, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
            'wql_set_tile_order': ([vp, C.c_int, C.c_int], C.c_int),
            'wql_get_tile_order': ([vp, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
            'wql_tile_variants': ([], C.c_int),
            'wql_use_graph': ([vp, C.c_int], C.c_int),
            'wql_profile': ([vp, C.c_int], C.c_int),
            'wql_volume_kernel_ms': ([vp], C.c_double),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


class WqlError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f'wql error {code}: {msg}')
        self.code = code


def check(rc):
    if rc != 0:
        raise WqlError(rc, lib().wql_last_error().decode())
    return rc
