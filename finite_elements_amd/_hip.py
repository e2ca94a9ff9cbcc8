"""ctypes binding of libfe_hip.so (the C ABI declared in include/fe_hip.h).

The library is built in-tree by ``finite_elements_amd.csrc.build`` (``__graft_entry__.build()``).
There is no CPU fallback: if the library is missing or cannot be loaded, every GPU entry point of this
package raises ``RuntimeError``.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'lib', 'libfe_hip.so')

FE_OK = 0
FE_ERR_ARG = -1
FE_ERR_NOT_CONVERGED = -2
FE_ERR_BREAKDOWN = -3
FE_ERR_UNSUPPORTED = -4
FE_ERR_NOMEM = -5

FE_KIND_TRI_ELASTIC = 1
FE_KIND_TET_ELASTIC = 2
FE_KIND_TRI_MAGNETIC = 3
FE_PLANE_STRAIN = 1
FE_PLANE_STRESS = 2
FE_STIFFNESS = 0
FE_MASS = 1
FE_LAYOUT_CSR = 0
FE_LAYOUT_BSR = 1
FE_LAYOUT_PLANES = 2
FE_PC_JACOBI = 1
FE_PC_BLOCK_JACOBI = 2
FE_COARSE_NOALLOC_MAX = 65536

FE_PAT_NNZ, FE_PAT_NNZB, FE_PAT_NINC, FE_PAT_NCONTRIB, FE_PAT_NBLK, FE_PAT_MAXINC, FE_PAT_NROWS = range(7)
FE_PAT_NINFO = 8
(FE_ARR_ROWPTR, FE_ARR_COLIND, FE_ARR_GATHER, FE_ARR_SEGPTR, FE_ARR_BPTR, FE_ARR_BCOL, FE_ARR_INCPTR,
 FE_ARR_INC, FE_ARR_BSEGPTR, FE_ARR_BSEG, FE_ARR_BLK) = range(11)

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double


class FeMesh(ctypes.Structure):
    _fields_ = [('kind', _i32), ('plane_mode', _i32), ('n_elem', _i64), ('n_nodes', _i64),
                ('x', _vp), ('y', _vp), ('z', _vp), ('conn', _vp), ('mat0', _vp), ('mat1', _vp),
                ('mat2', _vp), ('rho', _vp)]


class FeGraph(ctypes.Structure):
    _fields_ = [('n_nodes', _i64), ('nnzb', _i64), ('n_inc', _i64), ('n_blk', _i64),
                ('max_inc', _i32), ('threads', _i32), ('bptr', _vp), ('bcol', _vp), ('inc_ptr', _vp),
                ('inc', _vp), ('bseg_ptr', _vp), ('bseg', _vp), ('blk', _vp), ('max_elem', _i32),
                ('producers', _i32), ('welem_ptr', _vp), ('welem', _vp), ('inc_sa', _vp),
                ('bdiag', _vp), ('wg_hdr', _vp), ('welem_conn', _vp), ('n_welem', _i64),
                ('max_nodes', _i32), ('max_nbq', _i32), ('max_nseg', _i32), ('max_wnode', _i32),
                ('wnode', _vp), ('welem_local', _vp), ('wcode', _vp), ('brow', _vp),
                ('wmaps', _vp)]


class FeBsr(ctypes.Structure):
    _fields_ = [('n_brows', _i64), ('nnzb', _i64), ('dof', _i32), ('precision', _i32),
                ('layout', _i32), ('row_align', _i32), ('bptr', _vp), ('bcol', _vp), ('vals', _vp)]


class FePcgOpts(ctypes.Structure):
    _fields_ = [('tol', _dbl), ('max_iter', _i32), ('precond', _i32), ('mixed', _i32),
                ('check_every', _i32), ('ru_delta', _dbl), ('profile', _i32), ('graph', _i32)]


class FePcgStats(ctypes.Structure):
    _fields_ = [('iterations', _i32), ('reliable_updates', _i32), ('status', _i32), ('restarts', _i32),
                ('relres', _dbl), ('relres_rec', _dbl), ('bnorm', _dbl), ('spmv_ms', _dbl),
                ('spmv_samples', _i32), ('fallbacks', _i32)]


class FePcgSegStats(ctypes.Structure):
    _fields_ = [('iterations', _i32), ('reliable_updates', _i32), ('status', _i32), ('restarts', _i32),
                ('fallback', _i32), ('pad', _i32), ('relres', _dbl), ('relres_rec', _dbl), ('bnorm', _dbl)]


class FeCoarse(ctypes.Structure):
    _fields_ = [('n_nodes', _i64), ('n_agg', _i64), ('n_pairs', _i64), ('dof', _i32), ('nc', _i32),
                ('node_agg', _vp), ('agg_ptr', _vp), ('agg_nodes', _vp), ('pair_ptr', _vp),
                ('pair_ij', _vp), ('pair_q', _vp), ('pair_row', _vp), ('P', _vp), ('P32', _vp),
                ('xhat', _vp), ('xhat32', _vp), ('fixed_bits', _vp), ('T', _vp), ('T32', _vp),
                ('n_seg', _i32), ('max_block', _i32), ('blk_total', _i64), ('agg_seg', _vp),
                ('seg_of_agg', _vp), ('blk_ofs', _vp)]


class FeLobpcgOpts(ctypes.Structure):
    _fields_ = [('n_variants', _i32), ('block', _i32), ('n_wanted', _i32), ('largest', _i32),
                ('max_iter', _i32), ('check_every', _i32), ('tol', _dbl), ('k_stride', _i64),
                ('m_stride', _i64), ('t_stride', _i64)]


class FeLobpcgStats(ctypes.Structure):
    _fields_ = [('max_iterations', _i32), ('unconverged', _i32), ('failed', _i32), ('status', _i32)]


_P = ctypes.POINTER
# name -> (restype, argtypes); every function returns an int status unless noted.
SIGNATURES = {
    'fe_version': (_int, []),
    'fe_last_error': (ctypes.c_char_p, []),
    'fe_create': (_int, [_P(_vp), _int, _vp]),
    'fe_destroy': (_int, [_vp]),
    'fe_synchronize': (_int, [_vp]),
    'fe_ke_tri_elastic': (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _int, _i64, _int, _vp]),
    'fe_ke_tet_elastic': (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _int, _vp]),
    'fe_ke_tri_magnetic': (_int, [_vp, _vp, _vp, _vp, _vp, _i64, _int, _vp]),
    'fe_me_tri': (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _int, _vp]),
    'fe_me_tet': (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _int, _vp]),
    'fe_pattern_create': (_int, [_vp, _int, _int, _i64, _i64, _int, _int, _P(_vp)]),
    'fe_pattern_schedule': (_int, [_vp, _int]),
    'fe_pattern_info': (_int, [_vp, _P(_i64)]),
    'fe_pattern_copy': (_int, [_vp, _int, _vp]),
    'fe_pattern_destroy': (_int, [_vp]),
    'fe_assemble_gather': (_int, [_vp, _vp, _vp, _vp, _i64, _vp]),
    'fe_assemble_fused': (_int, [_vp, _P(FeMesh), _P(FeGraph), _int, _int, _vp]),
    'fe_ws_map_bytes': (_int, [_P(FeGraph), _P(_i64)]),
    'fe_ws_pack': (_int, [_vp, _P(FeGraph), _vp]),
    'fe_relayout': (_int, [_vp, _vp, _i64, _int, _int, _vp, _vp]),
    'fe_apply_dirichlet_bsr': (_int, [_vp, _P(FeBsr), _vp, _vp]),
    'fe_dirichlet_rhs': (_int, [_vp, _i64, _vp, _vp, _vp, _vp, _vp]),
    'fe_spmv': (_int, [_vp, _P(FeBsr), _vp, _vp]),
    'fe_spmv_csr': (_int, [_vp, _i64, _vp, _vp, _vp, _vp, _vp]),
    'fe_cast_f64_to_f32': (_int, [_vp, _i64, _vp, _vp]),
    'fe_bsr_convert': (_int, [_vp, _P(FeBsr), _P(FeBsr)]),
    'fe_element_fields': (_int, [_vp, _P(FeMesh), _vp, _vp, _vp, _vp]),
    'fe_schedule_elements': (_int, [_vp, _vp, _int, _i64, _vp, _i64, _vp, _vp, _vp, _P(_i32)]),
    'fe_schedule_ranges': (_int, [_vp, _i64, _int, _int, _vp, _P(_i64), _P(_i32)]),
    'fe_schedule_nodes': (_int, [_vp, _int, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _P(_i32)]),
    'fe_schedule_codes': (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i64, _vp, _vp]),
    'fe_bsr_repack': (_int, [_vp, _P(FeBsr), _P(FeBsr), _vp]),
    'fe_bsr_fold': (_int, [_vp, _P(FeBsr), _P(FeBsr), _vp, _vp, _vp]),
    'fe_precond_setup': (_int, [_vp, _P(FeBsr), _int, _vp]),
    'fe_pcg_workspace': (_int, [_vp, _P(FeBsr), _P(FePcgOpts), _P(_i64)]),
    'fe_pcg': (_int, [_vp, _P(FeBsr), _P(FeBsr), _vp, _vp, _vp, _vp, _vp, _i64, _P(FePcgOpts),
                      _P(FePcgStats)]),
    'fe_pcg_batched_workspace': (_int, [_vp, _P(FeBsr), _i32, _vp, _P(FePcgOpts), _P(_i64)]),
    'fe_pcg_batched': (_int, [_vp, _P(FeBsr), _P(FeBsr), _vp, _vp, _i32, _vp, _vp, _vp, _vp, _i64,
                              _P(FePcgOpts), _P(FePcgStats), _vp]),
    'fe_coarse_ld': (_i64, [_i64]),
    'fe_coarse_operator': (_int, [_vp, _P(FeBsr), _P(FeCoarse), _int, _vp, _vp, _vp]),
    'fe_coarse_apply': (_int, [_vp, _P(FeCoarse), _int, _vp, _vp, _vp, _vp]),
    'fe_coarse_apply_noalloc': (_int, [_vp, _P(FeCoarse), _int, _vp, _vp, _vp]),
    'fe_pcg_coarse_workspace': (_int, [_vp, _P(FeBsr), _P(FeCoarse), _P(FePcgOpts), _P(_i64)]),
    'fe_pcg_coarse': (_int, [_vp, _P(FeBsr), _P(FeBsr), _vp, _vp, _P(FeCoarse), _vp, _vp, _vp, _vp,
                             _vp, _i64, _P(FePcgOpts), _P(FePcgStats)]),
    'fe_pcg_batched_coarse_workspace': (_int, [_vp, _P(FeBsr), _i32, _vp, _P(FeCoarse), _P(FePcgOpts),
                                               _P(_i64)]),
    'fe_pcg_batched_coarse': (_int, [_vp, _P(FeBsr), _P(FeBsr), _vp, _vp, _i32, _vp, _P(FeCoarse), _vp, _vp,
                                     _vp, _vp, _vp, _i64, _P(FePcgOpts), _P(FePcgStats), _vp]),
    'fe_lobpcg_workspace': (_int, [_vp, _P(FeBsr), _P(FeLobpcgOpts), _P(_i64)]),
    'fe_lobpcg': (_int, [_vp, _P(FeBsr), _P(FeBsr), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64,
                         _P(FeLobpcgOpts), _P(FeLobpcgStats)]),
    'fe_lobpcg_coarse_workspace': (_int, [_vp, _P(FeBsr), _P(FeCoarse), _P(FeLobpcgOpts), _P(_i64)]),
    'fe_lobpcg_coarse': (_int, [_vp, _P(FeBsr), _P(FeBsr), _vp, _vp, _P(FeCoarse), _vp, _i64, _vp, _vp,
                                _vp, _vp, _vp, _vp, _i64, _P(FeLobpcgOpts), _P(FeLobpcgStats)]),
    'fe_sym_eig': (_int, [_vp, _int, _int, _vp, _vp, _vp]),
    'fe_gnn_num_params': (_int, [_int]),
    'fe_gnn_features': (_int, [_vp, _P(FeBsr), _vp, _vp, _vp]),
    'fe_gnn_forward': (_int, [_vp, _i64, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    'fe_gnn_loss_workspace': (_int, [_i64, _P(_i64)]),
    'fe_gnn_loss_grad': (_int, [_vp, _P(FeBsr), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _int]),
    'fe_gnn_backward_workspace': (_int, [_i64, _P(_i64)]),
    'fe_gnn_backward': (_int, [_vp, _i64, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    'fe_gnn_adam': (_int, [_vp, _int, _vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                           ctypes.c_float, _int, ctypes.c_float]),
}

_lib = None


class FeError(RuntimeError):
    """A libfe_hip call failed."""

    def __init__(self, name, code, message):
        super().__init__(f'{name} failed with status {code}: {message}')
        self.code = code


def load():
    """Load libfe_hip.so (raises RuntimeError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f'{LIB_PATH} is not built; run finite_elements_amd.csrc.build.build() '
                           '(or __graft_entry__.build()) first')
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def call(name, *args):
    """Call ``name``; raise ValueError for argument errors and FeError for everything else."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != FE_OK:
        msg = lib.fe_last_error().decode(errors='replace')
        if rc == FE_ERR_ARG:
            raise ValueError(f'{name}: {msg}')
        raise FeError(name, rc, msg)
    return rc
