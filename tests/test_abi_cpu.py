"""The C ABI: libfe_hip.so loads, exports every entry point include/fe_hip.h declares, its struct
layouts match the ctypes mirrors, and argument errors come back as status codes (no GPU needed)."""

import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from finite_elements_amd import _hip
from tests.conftest import ROOT

HEADER = os.path.join(ROOT, 'include', 'fe_hip.h')


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r'^\s*(?:int|int64_t|const char\*)\s+(fe_\w+)\s*\(', text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _hip.load()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
        assert name in _hip.SIGNATURES, f'{name} has no ctypes signature'
    assert set(_hip.SIGNATURES) == set(names)
    assert lib.fe_version() == 1


def test_struct_layouts_match_header():
    structs = {'fe_mesh': _hip.FeMesh, 'fe_graph': _hip.FeGraph, 'fe_bsr': _hip.FeBsr,
               'fe_pcg_opts': _hip.FePcgOpts, 'fe_pcg_stats': _hip.FePcgStats,
               'fe_pcg_seg_stats': _hip.FePcgSegStats, 'fe_coarse': _hip.FeCoarse,
               'fe_lobpcg_opts': _hip.FeLobpcgOpts, 'fe_lobpcg_stats': _hip.FeLobpcgStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fe_hip.h"', 'int main(void) {']
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append('return 0; }')
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, 'layout.c')
        exe = os.path.join(d, 'layout')
        open(src, 'w').write('\n'.join(lines))
        subprocess.run(['gcc', '-I', os.path.join(ROOT, 'include'), src, '-o', exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    got = dict(line.rsplit(' ', 1) for line in out.strip().splitlines())
    for cname, cls in structs.items():
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f'{cname}.{fname}']) == getattr(cls, fname).offset, f'{cname}.{fname}'


def test_argument_errors_are_status_codes():
    lib = _hip.load()
    p = ctypes.c_void_p()
    conn = np.zeros((3, 1), dtype=np.int32)
    rc = lib.fe_pattern_create(conn.ctypes.data_as(ctypes.c_void_p), 7, 2, 1, 3, 1, 0, ctypes.byref(p))
    assert rc == _hip.FE_ERR_ARG
    assert b'npe' in lib.fe_last_error()
    with pytest.raises(ValueError):
        _hip.call('fe_pattern_create', conn.ctypes.data_as(ctypes.c_void_p), 3, 2, 1, 3, 1, 0, None)
    bad = np.array([[0], [1], [9]], dtype=np.int32)  # node index out of range
    with pytest.raises(ValueError, match='out of range'):
        _hip.call('fe_pattern_create', bad.ctypes.data_as(ctypes.c_void_p), 3, 2, 1, 3, 1, 0,
                  ctypes.byref(p))
    # null handle on a device entry point is an argument error, not a crash
    assert lib.fe_synchronize(None) == _hip.FE_ERR_ARG


def test_no_cpu_fallback_without_gpu():
    torch = pytest.importorskip('torch')
    if torch.cuda.is_available():
        pytest.skip('a GPU is present')
    from finite_elements_amd import device
    with pytest.raises(RuntimeError, match='no ROCm GPU'):
        device.Context.get(0)
