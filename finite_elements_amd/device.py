"""Device plumbing: torch tensors on the ROCm device handed to libfe_hip.so through its C ABI.

PyTorch only owns device memory and the stream here; every computation on the path (element matrices,
assembly, Dirichlet elimination, SpMV, preconditioner setup, PCG) runs in the hand-written HIP kernels
of libfe_hip.so.  There is no CPU fallback: without a ROCm device or without the built library these
functions raise ``RuntimeError``.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _hip

KIND_INFO = {
    # kind -> (C kind id, nodes per element, dof per node)
    'tri_elastic': (_hip.FE_KIND_TRI_ELASTIC, 3, 2),
    'tet_elastic': (_hip.FE_KIND_TET_ELASTIC, 4, 3),
    'tri_magnetic': (_hip.FE_KIND_TRI_MAGNETIC, 3, 1),
}


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError('finite_elements_amd: no ROCm GPU is visible; the HIP path has no CPU '
                           'fallback')


class Context:
    """A libfe_hip handle bound to torch's current stream on one device."""

    _by_device = {}

    def __init__(self, device: int = 0, stream: Optional[torch.cuda.Stream] = None):
        require_gpu()
        _hip.load()
        self.device = device
        self.torch_device = torch.device('cuda', device)
        with torch.cuda.device(device):
            self.stream = torch.cuda.current_stream(device) if stream is None else stream
        h = ctypes.c_void_p()
        _hip.call('fe_create', ctypes.byref(h), device, ctypes.c_void_p(self.stream.cuda_stream))
        self.h = h
        self._sides = {}

    def side(self, index: int = 0) -> 'Context':
        """Another handle on a stream of its own (default priority), for work that may run beside
        the main stream's; ``index`` names one of several, each created once."""
        ctx = self._sides.get(index)
        if ctx is None:
            with torch.cuda.device(self.device):
                ctx = Context(self.device, torch.cuda.Stream(self.device))
            self._sides[index] = ctx
        return ctx

    @classmethod
    def get(cls, device: int = 0) -> 'Context':
        ctx = cls._by_device.get(device)
        if ctx is None:
            ctx = cls(device)
            cls._by_device[device] = ctx
        return ctx

    def synchronize(self):
        _hip.call('fe_synchronize', self.h)

    def tensor(self, arr, dtype=None):
        t = torch.as_tensor(np.ascontiguousarray(arr))
        if dtype is not None:
            t = t.to(dtype)
        return t.to(self.torch_device)

    def empty(self, n, dtype=torch.float64):
        return torch.empty(int(n), dtype=dtype, device=self.torch_device)

    def zeros(self, n, dtype=torch.float64):
        return torch.zeros(int(n), dtype=dtype, device=self.torch_device)


# ------------------------------------------------------------------------------------------------
# pattern (host C++ builder)
# ------------------------------------------------------------------------------------------------
class Pattern:
    """Node graph, scalar CSR pattern and contribution maps of a mesh (built by fe_pattern_create).

    ``conn`` is (E, npe) on the host.  ``with_gather`` also builds the per-entry contribution map of
    the two-phase assembly (m*m int32 per element).
    """

    def __init__(self, conn: np.ndarray, dof: int, n_nodes: int, sym: bool = True,
                 with_gather: bool = False, max_inc: Optional[int] = None):
        _hip.load()
        conn = np.ascontiguousarray(np.asarray(conn, dtype=np.int32))
        self.n_elem, self.npe = conn.shape
        self.dof = dof
        self.n_nodes = int(n_nodes)
        self.sym = bool(sym)
        soa = np.ascontiguousarray(conn.T)
        p = ctypes.c_void_p()
        _hip.call('fe_pattern_create', soa.ctypes.data_as(ctypes.c_void_p), self.npe, dof,
                  self.n_elem, self.n_nodes, int(sym), int(with_gather), ctypes.byref(p))
        self._p = p
        try:
            if max_inc is not None:
                _hip.call('fe_pattern_schedule', p, int(max_inc))
            self._export(with_gather)
        finally:
            _hip.call('fe_pattern_destroy', p)
            self._p = None

    def _info(self):
        info = (ctypes.c_int64 * _hip.FE_PAT_NINFO)()
        _hip.call('fe_pattern_info', self._p, info)
        return list(info)

    def _copy(self, which, n, dtype):
        out = np.empty(int(n), dtype=dtype)
        if n:
            _hip.call('fe_pattern_copy', self._p, which, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def _export(self, with_gather):
        info = self._info()
        self.nnz = info[_hip.FE_PAT_NNZ]
        self.nnzb = info[_hip.FE_PAT_NNZB]
        self.n_inc = info[_hip.FE_PAT_NINC]
        self.n_rows = info[_hip.FE_PAT_NROWS]
        self.rowptr = self._copy(_hip.FE_ARR_ROWPTR, self.n_rows + 1, np.int32)
        self.colind = self._copy(_hip.FE_ARR_COLIND, self.nnz, np.int32)
        self.bptr = self._copy(_hip.FE_ARR_BPTR, self.n_nodes + 1, np.int32)
        self.bcol = self._copy(_hip.FE_ARR_BCOL, self.nnzb, np.int32)
        self.inc_ptr = self._copy(_hip.FE_ARR_INCPTR, self.n_nodes + 1, np.int32)
        self.inc = self._copy(_hip.FE_ARR_INC, self.n_inc, np.int32)
        self.bseg_ptr = self._copy(_hip.FE_ARR_BSEGPTR, self.nnzb + 1, np.int32)
        self.bseg = self._copy(_hip.FE_ARR_BSEG, self.n_inc * self.npe, np.uint16)
        self._set_schedule(info)
        if with_gather:
            ncontrib = info[_hip.FE_PAT_NCONTRIB]
            self.gather = self._copy(_hip.FE_ARR_GATHER, ncontrib, np.int32)
            self.seg_ptr = self._copy(_hip.FE_ARR_SEGPTR, self.nnz + 1, np.int32)
        else:
            self.gather = None
            self.seg_ptr = None

    def _set_schedule(self, info):
        self.n_blk = info[_hip.FE_PAT_NBLK]
        self.max_inc = info[_hip.FE_PAT_MAXINC]
        self.blk = self._copy(_hip.FE_ARR_BLK, self.n_blk + 1, np.int32)

    @staticmethod
    def schedule_for(inc_ptr: np.ndarray, max_inc: int, max_nodes: int = 256):
        """Row-block schedule (the greedy rule of fe_pattern_schedule, fe_schedule_ranges) for a
        different max_inc: (blk, largest incidence count of a range)."""
        inc_ptr = np.ascontiguousarray(np.asarray(inc_ptr, dtype=np.int32))
        n = inc_ptr.size - 1
        counts = np.diff(inc_ptr)
        if counts.size and counts.max() > max_inc:
            raise ValueError(f'a node has {counts.max()} incidences > max_inc {max_inc}')
        blk = np.empty(n + 1, dtype=np.int32)
        n_blk = ctypes.c_int64()
        biggest = ctypes.c_int32()
        _hip.call('fe_schedule_ranges', inc_ptr.ctypes.data_as(ctypes.c_void_p), n, int(max_inc),
                  int(max_nodes), blk.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n_blk),
                  ctypes.byref(biggest))
        return blk[:n_blk.value + 1].copy(), int(biggest.value)


# ------------------------------------------------------------------------------------------------
# device mesh / graph / matrices
# ------------------------------------------------------------------------------------------------
class DeviceMesh:
    """SoA mesh on the device: coordinates, connectivity [npe][E] and per-element materials."""

    def __init__(self, ctx: Context, kind: str, coords: np.ndarray, conn: np.ndarray,
                 materials: dict, plane_mode: int = 0):
        self.ctx = ctx
        self.kind = kind
        self.kind_id, self.npe, self.dof = KIND_INFO[kind]
        coords = np.asarray(coords, dtype=np.float64)
        conn = np.asarray(conn, dtype=np.int32)
        self.n_nodes = coords.shape[0]
        self.n_elem = conn.shape[0]
        self.plane_mode = plane_mode
        self.x = ctx.tensor(coords[:, 0])
        self.y = ctx.tensor(coords[:, 1])
        self.z = ctx.tensor(coords[:, 2]) if coords.shape[1] > 2 else None
        self.conn = ctx.tensor(np.ascontiguousarray(conn.T))

        def mat(key):
            v = materials.get(key)
            if v is None:
                return None
            v = np.asarray(v, dtype=np.float64)
            if v.ndim == 0:
                v = np.full(self.n_elem, float(v))
            return ctx.tensor(v)

        if kind == 'tri_elastic':
            self.mat0, self.mat1, self.mat2 = mat('E'), mat('nu'), mat('thickness')
        elif kind == 'tet_elastic':
            self.mat0, self.mat1, self.mat2 = mat('E'), mat('nu'), None
        else:
            self.mat0, self.mat1, self.mat2 = mat('mu'), None, None
        self.rho = mat('rho')

    def struct(self) -> _hip.FeMesh:
        return _hip.FeMesh(self.kind_id, self.plane_mode, self.n_elem, self.n_nodes, _vp(self.x),
                           _vp(self.y), _vp(self.z), _vp(self.conn), _vp(self.mat0), _vp(self.mat1),
                           _vp(self.mat2), _vp(self.rho))


def element_schedule(pat: Pattern, blk: np.ndarray):
    """(welem_ptr, welem, inc_sa, max_elem) of the fused assembly for the node ranges ``blk``
    (fe_schedule_elements, host code): the ascending distinct elements of every workgroup and, per
    incidence, (slot << 2) | local vertex."""
    blk = np.ascontiguousarray(np.asarray(blk, dtype=np.int32))
    n_inc = int(pat.inc.size)
    wptr = np.empty(blk.size, dtype=np.int32)
    welem = np.empty(max(n_inc, 1), dtype=np.int32)
    inc_sa = np.empty(max(n_inc, 1), dtype=np.uint16)
    max_elem = ctypes.c_int32()
    _hip.call('fe_schedule_elements', pat.inc_ptr.ctypes.data_as(ctypes.c_void_p),
              pat.inc.ctypes.data_as(ctypes.c_void_p), pat.npe, pat.n_nodes,
              blk.ctypes.data_as(ctypes.c_void_p), blk.size - 1,
              wptr.ctypes.data_as(ctypes.c_void_p), welem.ctypes.data_as(ctypes.c_void_p),
              inc_sa.ctypes.data_as(ctypes.c_void_p), ctypes.byref(max_elem))
    return wptr, welem[:int(wptr[-1])], inc_sa[:n_inc], int(max_elem.value)


class DeviceGraph:
    """Pattern arrays on the device (for the fused assembly, SpMV and the solver)."""

    def __init__(self, ctx: Context, pat: Pattern):
        self.ctx = ctx
        self.pattern = pat
        self.n_nodes = pat.n_nodes
        self.nnzb = pat.nnzb
        self.dof = pat.dof
        self.bptr = ctx.tensor(pat.bptr)
        self.bcol = ctx.tensor(pat.bcol)
        self.inc_ptr = ctx.tensor(pat.inc_ptr)
        self.inc = ctx.tensor(pat.inc)
        self.bseg_ptr = ctx.tensor(pat.bseg_ptr)
        self.bseg = ctx.tensor(pat.bseg.view(np.int16))
        rows = np.repeat(np.arange(self.n_nodes, dtype=np.int32), np.diff(pat.bptr))
        bdiag = np.full(max(self.n_nodes, 1), -1, dtype=np.int32)
        q = np.nonzero(pat.bcol == rows)[0]
        bdiag[rows[q]] = q
        self.bdiag = ctx.tensor(bdiag)
        self._conn = None
        self.set_schedule(pat.blk, pat.max_inc)
        self._padded = None

    def attach_connectivity(self, conn: np.ndarray):
        """Host connectivity (E, npe) of the mesh the schedule serves: every schedule then carries
        the node ids of its element lists (one load level less in the fused assembly)."""
        self._conn = np.ascontiguousarray(np.asarray(conn, dtype=np.int32))
        self._set_welem_conn()

    def _set_welem_conn(self):
        self.wnode = self.welem_local = self.wcode = self.brow = self.wmaps = None
        self.max_wnode = 0
        if self._conn is None or self._welem_host.size == 0:
            self.welem_conn = None
            return
        wconn = np.ascontiguousarray(self._conn[self._welem_host].T)
        self.welem_conn = self.ctx.tensor(wconn)
        if self.producers > 0:
            self._set_ws_lists(wconn)

    def _set_ws_lists(self, wconn):
        """The wave-specialised kernel's node lists and resolved block maps for this schedule."""
        # node lists of the element lists (fe_schedule_nodes) and their slices in header words 10, 11
        npe, n_welem = wconn.shape
        wptr = self._welem_ptr_host
        wnode_ptr = np.empty(self.n_blk + 1, dtype=np.int32)
        wnode = np.empty(max(npe * n_welem, 1), dtype=np.int32)
        local = np.empty((n_welem, 4), dtype=np.uint16)
        mx = ctypes.c_int32()
        _hip.call('fe_schedule_nodes', wconn.ctypes.data_as(ctypes.c_void_p), npe, n_welem,
                  wptr.ctypes.data_as(ctypes.c_void_p), self.n_blk, self.n_nodes,
                  wnode_ptr.ctypes.data_as(ctypes.c_void_p), wnode.ctypes.data_as(ctypes.c_void_p),
                  local.ctypes.data_as(ctypes.c_void_p), ctypes.byref(mx))
        self.max_wnode = int(mx.value)
        self.wnode = self.ctx.tensor(wnode[:int(wnode_ptr[-1])])
        self.welem_local = self.ctx.tensor(local.view(np.int16))
        hdr = self._hdr_host.copy()
        hdr[:, 10], hdr[:, 11] = wnode_ptr[:-1], np.diff(wnode_ptr)
        self.wg_hdr = self.ctx.tensor(hdr.reshape(-1))
        # the block maps resolved against this schedule's element lists (fe_schedule_codes)
        p = self.pattern
        code = np.empty(max(p.bseg.size, 1), dtype=np.uint16)
        brow = np.empty(max(p.nnzb, 1), dtype=np.uint16)
        blk = self._blk_host
        _hip.call('fe_schedule_codes', p.bptr.ctypes.data_as(ctypes.c_void_p),
                  p.bcol.ctypes.data_as(ctypes.c_void_p),
                  p.bseg_ptr.ctypes.data_as(ctypes.c_void_p), p.bseg.ctypes.data_as(ctypes.c_void_p),
                  p.inc_ptr.ctypes.data_as(ctypes.c_void_p),
                  self._inc_sa_host.ctypes.data_as(ctypes.c_void_p), p.n_nodes,
                  blk.ctypes.data_as(ctypes.c_void_p), self.n_blk,
                  code.ctypes.data_as(ctypes.c_void_p), brow.ctypes.data_as(ctypes.c_void_p))
        self.wcode = self.ctx.tensor(code.view(np.int16))
        self.brow = self.ctx.tensor(brow.view(np.int16))

    def padded(self) -> 'PaddedGraph':
        """The padded pattern of the same rows (row_align = 4), built once."""
        if self._padded is None:
            self._padded = PaddedGraph(self)
        return self._padded

    def set_schedule(self, blk: np.ndarray, max_inc: int, threads: int = 256, producers: int = 0):
        """Row-block schedule of the fused assembly and its element schedule (the distinct elements
        of every workgroup and each incidence's slot among them).  ``producers`` > 0 selects the
        wave-specialised persistent kernel with that many producer wavefronts."""
        blk = np.ascontiguousarray(np.asarray(blk, dtype=np.int32))
        wptr, welem, inc_sa, max_elem = element_schedule(self.pattern, blk)
        self.blk = self.ctx.tensor(blk)
        self.n_blk = blk.size - 1
        self.max_inc = int(max_inc)
        self.threads = int(threads)
        self.max_elem = max_elem
        self.welem_ptr = self.ctx.tensor(wptr)
        self.welem = self.ctx.tensor(welem if welem.size else np.zeros(1, np.int32))
        self.inc_sa = self.ctx.tensor((inc_sa if inc_sa.size else np.zeros(1, np.uint16)).view(np.int16))
        # per-workgroup header: every index the workgroup starts from, in one 48-byte record
        p = self.pattern
        n0, n1 = blk[:-1].astype(np.int64), blk[1:].astype(np.int64)
        hdr = np.zeros((self.n_blk, 12), dtype=np.int32)
        if self.n_blk:
            B0, B1 = p.bptr[n0], p.bptr[n1]
            hdr[:, 0], hdr[:, 1] = n0, n1
            hdr[:, 2], hdr[:, 3] = wptr[:-1], np.diff(wptr)
            hdr[:, 4], hdr[:, 5] = p.inc_ptr[n0], p.inc_ptr[n1] - p.inc_ptr[n0]
            hdr[:, 6], hdr[:, 7] = B0, B1 - B0
            hdr[:, 8], hdr[:, 9] = p.bseg_ptr[B0], p.bseg_ptr[B1] - p.bseg_ptr[B0]
        self.wg_hdr = self.ctx.tensor(hdr.reshape(-1) if hdr.size else np.zeros(12, np.int32))
        self._hdr_host = hdr
        self._welem_ptr_host = wptr
        self._blk_host = blk
        self._inc_sa_host = np.ascontiguousarray(inc_sa if inc_sa.size else np.zeros(1, np.uint16))
        self.max_nodes = int(np.diff(blk).max()) if self.n_blk else 0
        self.max_nbq = int(hdr[:, 7].max()) if self.n_blk else 0
        self.max_nseg = int(hdr[:, 9].max()) if self.n_blk else 0
        self.producers = int(producers)
        self._welem_host = welem
        self._set_welem_conn()

    def struct(self) -> _hip.FeGraph:
        return _hip.FeGraph(self.n_nodes, self.nnzb, int(self.inc.numel()), self.n_blk, self.max_inc,
                            self.threads,
                            _vp(self.bptr), _vp(self.bcol), _vp(self.inc_ptr), _vp(self.inc),
                            _vp(self.bseg_ptr), _vp(self.bseg), _vp(self.blk), self.max_elem,
                            self.producers, _vp(self.welem_ptr), _vp(self.welem), _vp(self.inc_sa),
                            _vp(self.bdiag), _vp(self.wg_hdr), _vp(self.welem_conn),
                            int(self._welem_host.size), self.max_nodes, self.max_nbq, self.max_nseg,
                            self.max_wnode, _vp(self.wnode), _vp(self.welem_local), _vp(self.wcode),
                            _vp(self.brow), _vp(self.wmaps))

    def ws_maps(self):
        """The map images of the wave-specialised assembly for this schedule (fe_ws_pack), built on
        the device on first use."""
        if self.wmaps is None and self.wcode is not None:
            nb = ctypes.c_int64()
            gs = self.struct()
            _hip.call('fe_ws_map_bytes', ctypes.byref(gs), ctypes.byref(nb))
            n4 = max(self.n_blk * nb.value // 16, 1)
            with torch.cuda.stream(self.ctx.stream):  # zeroed on the stream fe_ws_pack runs on
                buf = torch.zeros((n4, 4), dtype=torch.int32, device=self.ctx.torch_device)
            _hip.call('fe_ws_pack', self.ctx.h, ctypes.byref(gs), _vp(buf))
            self.wmaps = buf
        return self.wmaps


def pad_pattern(bptr: np.ndarray, bcol: np.ndarray, align: int = 4):
    """(bptr, bcol, src_block) of the padded pattern: every block row extended to a multiple of
    ``align`` blocks by blocks whose column is the row itself; src_block[q] is the original block of
    padded block q, or -1 for padding."""
    bptr = np.asarray(bptr, dtype=np.int64)
    n = bptr.size - 1
    nnzb = int(bptr[-1])
    deg = np.diff(bptr)
    pdeg = (deg + align - 1) // align * align
    pptr = np.zeros(n + 1, dtype=np.int64)
    pptr[1:] = np.cumsum(pdeg)
    if pptr[-1] >= 2 ** 31:
        raise ValueError('padded pattern exceeds 2^31 blocks')
    pos = np.repeat(pptr[:-1] - bptr[:-1], deg) + np.arange(nnzb, dtype=np.int64)
    pcol = np.repeat(np.arange(n, dtype=np.int32), pdeg)
    pcol[pos] = np.asarray(bcol, dtype=np.int32)
    src = np.full(int(pptr[-1]), -1, dtype=np.int32)
    src[pos] = np.arange(nnzb, dtype=np.int32)
    return pptr.astype(np.int32), pcol, src


class PaddedGraph:
    """Padded pattern of a DeviceGraph's rows: each block row is extended to a multiple of 4 blocks
    with zero blocks whose column is the row itself, so that the solvers' SpMV reads 4 (FP32) or 2
    (FP64) blocks of a value plane with one 16-byte load.  ``src_block`` maps every padded block to
    its block in the original pattern (-1 for padding)."""

    row_align = 4

    def __init__(self, graph: DeviceGraph):
        a = self.row_align
        self.ctx = graph.ctx
        self.base = graph
        self.n_nodes = graph.n_nodes
        pptr, bcol, src = pad_pattern(graph.pattern.bptr, graph.pattern.bcol, a)
        self.nnzb = int(pptr[-1])
        self.bptr = self.ctx.tensor(pptr)
        self.bcol = self.ctx.tensor(bcol)
        self.src_block = self.ctx.tensor(src)


@dataclass
class BsrMatrix:
    """Node-block matrix on the device: ``vals`` holds nnzb * dof * dof values in ``layout``
    (FE_LAYOUT_BSR: [q][di][dj]; FE_LAYOUT_PLANES: [di*dof+dj][q], the solver's layout)."""
    graph: DeviceGraph
    vals: torch.Tensor
    dof: int
    layout: int = _hip.FE_LAYOUT_BSR

    @property
    def precision(self):
        return 0 if self.vals.dtype == torch.float64 else 1

    @property
    def n(self):
        return self.graph.n_nodes * self.dof

    def struct(self) -> _hip.FeBsr:
        return _hip.FeBsr(self.graph.n_nodes, self.graph.nnzb, self.dof, self.precision, self.layout,
                          getattr(self.graph, 'row_align', 0), _vp(self.graph.bptr),
                          _vp(self.graph.bcol), _vp(self.vals))

    def padded(self, dtype=torch.float64) -> 'BsrMatrix':
        """Iteration copy on the padded pattern (FE_LAYOUT_PLANES, row_align 4) in ``dtype``."""
        pg = self.graph.padded()
        out = BsrMatrix(pg, torch.empty(pg.nnzb * self.dof * self.dof, dtype=dtype,
                                        device=self.vals.device), self.dof, _hip.FE_LAYOUT_PLANES)
        a, b = self.struct(), out.struct()
        _hip.call('fe_bsr_repack', self.graph.ctx.h, ctypes.byref(a), ctypes.byref(b),
                  _vp(pg.src_block))
        return out

    def convert(self, layout: int, dtype=torch.float64) -> 'BsrMatrix':
        out = BsrMatrix(self.graph, torch.empty(self.vals.numel(), dtype=dtype, device=self.vals.device),
                        self.dof, layout)
        a, b = self.struct(), out.struct()
        _hip.call('fe_bsr_convert', self.graph.ctx.h, ctypes.byref(a), ctypes.byref(b))
        return out

    def to_f32(self) -> 'BsrMatrix':
        return self.convert(self.layout, torch.float32)

    def clone(self) -> 'BsrMatrix':
        return BsrMatrix(self.graph, self.vals.clone(), self.dof, self.layout)

    def csr_values(self) -> torch.Tensor:
        """Values permuted to the scalar-CSR layout of the pattern (device tensor, FP64)."""
        src = self if (self.layout == _hip.FE_LAYOUT_BSR and self.precision == 0) else \
            self.convert(_hip.FE_LAYOUT_BSR, torch.float64)
        out = torch.empty_like(src.vals)
        relayout(self.graph.ctx, self.graph, self.dof, _hip.FE_LAYOUT_CSR, src.vals, out)
        return out


# ------------------------------------------------------------------------------------------------
# kernels
# ------------------------------------------------------------------------------------------------
def element_fields(mesh: DeviceMesh, u: torch.Tensor):
    """Per-element fields of the solution u (device FP64, interleaved DOFs): elasticity ->
    (strain (ns, E), stress (ns, E), energy (E,)) with ns = 3 (tri) or 6 (tet), Voigt order;
    magnetic -> (B (2, E), None, None)."""
    ctx = mesh.ctx
    E = mesh.n_elem
    if u.dtype != torch.float64 or u.numel() != mesh.n_nodes * mesh.dof:
        raise ValueError('u must be the FP64 solution vector of this mesh')
    u = u.contiguous()
    dev = ctx.torch_device
    if mesh.kind == 'tri_magnetic':
        b = torch.empty((2, E), dtype=torch.float64, device=dev)
        _hip.call('fe_element_fields', ctx.h, ctypes.byref(mesh.struct()), _vp(u), _vp(b), None, None)
        return b, None, None
    ns = 3 if mesh.kind == 'tri_elastic' else 6
    strain = torch.empty((ns, E), dtype=torch.float64, device=dev)
    stress = torch.empty((ns, E), dtype=torch.float64, device=dev)
    energy = torch.empty(E, dtype=torch.float64, device=dev)
    _hip.call('fe_element_fields', ctx.h, ctypes.byref(mesh.struct()), _vp(u), _vp(strain),
              _vp(stress), _vp(energy))
    return strain, stress, energy


def element_matrices(mesh: DeviceMesh, what: str = 'stiffness', sym: bool = False) -> torch.Tensor:
    """Entry-major element matrices ke[k, e] (k over m*m, or the m(m+1)/2 upper triangle)."""
    ctx = mesh.ctx
    m = mesh.npe * mesh.dof
    rows = m * (m + 1) // 2 if sym else m * m
    ke = torch.empty((rows, mesh.n_elem), dtype=torch.float64, device=ctx.torch_device)
    if mesh.n_elem == 0:
        return ke
    h, s = ctx.h, int(sym)
    if what == 'stiffness':
        if mesh.kind == 'tri_elastic':
            _hip.call('fe_ke_tri_elastic', h, _vp(mesh.x), _vp(mesh.y), _vp(mesh.conn), _vp(mesh.mat0),
                      _vp(mesh.mat1), _vp(mesh.mat2), mesh.plane_mode, mesh.n_elem, s, _vp(ke))
        elif mesh.kind == 'tet_elastic':
            _hip.call('fe_ke_tet_elastic', h, _vp(mesh.x), _vp(mesh.y), _vp(mesh.z), _vp(mesh.conn),
                      _vp(mesh.mat0), _vp(mesh.mat1), mesh.n_elem, s, _vp(ke))
        else:
            _hip.call('fe_ke_tri_magnetic', h, _vp(mesh.x), _vp(mesh.y), _vp(mesh.conn),
                      _vp(mesh.mat0), mesh.n_elem, s, _vp(ke))
    elif what == 'mass':
        if mesh.rho is None:
            raise ValueError('mass matrices need a mass density')
        if mesh.kind == 'tri_elastic':
            _hip.call('fe_me_tri', h, _vp(mesh.x), _vp(mesh.y), _vp(mesh.conn), _vp(mesh.rho),
                      _vp(mesh.mat2), mesh.n_elem, s, _vp(ke))
        elif mesh.kind == 'tet_elastic':
            _hip.call('fe_me_tet', h, _vp(mesh.x), _vp(mesh.y), _vp(mesh.z), _vp(mesh.conn),
                      _vp(mesh.rho), mesh.n_elem, s, _vp(ke))
        else:
            raise AttributeError("'MagneticElement2D' object has no attribute 'elementary_mass_matrix'")
    else:
        raise ValueError(f'unknown matrix {what!r}')
    return ke


def assemble_gather(ctx: Context, ke: torch.Tensor, gather: torch.Tensor, seg_ptr: torch.Tensor,
                    nnz: int) -> torch.Tensor:
    vals = ctx.empty(nnz)
    _hip.call('fe_assemble_gather', ctx.h, _vp(ke), _vp(gather), _vp(seg_ptr), int(nnz), _vp(vals))
    return vals


def assemble_fused(mesh: DeviceMesh, graph: DeviceGraph, what: str = 'stiffness',
                   layout: int = _hip.FE_LAYOUT_BSR, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    ctx = mesh.ctx
    n = graph.nnzb * mesh.dof * mesh.dof
    vals = ctx.empty(n) if out is None else out
    w = _hip.FE_STIFFNESS if what == 'stiffness' else _hip.FE_MASS
    if graph.producers > 0:
        graph.ws_maps()
    ms, gs = mesh.struct(), graph.struct()
    _hip.call('fe_assemble_fused', ctx.h, ctypes.byref(ms), ctypes.byref(gs), w, layout, _vp(vals))
    return vals


def relayout(ctx: Context, graph: DeviceGraph, dof: int, to_layout: int, src: torch.Tensor,
             dst: torch.Tensor):
    _hip.call('fe_relayout', ctx.h, _vp(graph.bptr), graph.n_nodes, dof, to_layout, _vp(src), _vp(dst))


def cast_f64_to_f32(ctx: Context, src: torch.Tensor, dst: torch.Tensor):
    _hip.call('fe_cast_f64_to_f32', ctx.h, src.numel(), _vp(src), _vp(dst))


def spmv(A: BsrMatrix, x: torch.Tensor, y: Optional[torch.Tensor] = None) -> torch.Tensor:
    ctx = A.graph.ctx
    if y is None:
        y = torch.empty_like(x)
    s = A.struct()
    _hip.call('fe_spmv', ctx.h, ctypes.byref(s), _vp(x), _vp(y))
    return y


def spmv_csr(ctx: Context, rowptr, colind, vals, x, y=None):
    if y is None:
        y = torch.empty_like(x)
    _hip.call('fe_spmv_csr', ctx.h, rowptr.numel() - 1, _vp(rowptr), _vp(colind), _vp(vals), _vp(x),
              _vp(y))
    return y


def apply_dirichlet(A: BsrMatrix, fixed: torch.Tensor) -> BsrMatrix:
    """Copy of A with rows/columns of fixed DOFs replaced by identity."""
    out = A.clone()
    s = A.struct()
    _hip.call('fe_apply_dirichlet_bsr', A.graph.ctx.h, ctypes.byref(s), _vp(fixed), _vp(out.vals))
    return out


def dirichlet_rhs(ctx: Context, fixed, g, f, kg):
    b = torch.empty_like(f)
    _hip.call('fe_dirichlet_rhs', ctx.h, f.numel(), _vp(fixed), _vp(g), _vp(f), _vp(kg), _vp(b))
    return b


def precond_setup(A: BsrMatrix, kind: str = 'block_jacobi') -> torch.Tensor:
    ctx = A.graph.ctx
    k = _hip.FE_PC_JACOBI if kind == 'jacobi' else _hip.FE_PC_BLOCK_JACOBI
    size = A.n if kind == 'jacobi' else A.graph.n_nodes * A.dof * A.dof
    out = ctx.empty(size)
    s = A.struct()
    _hip.call('fe_precond_setup', ctx.h, ctypes.byref(s), k, _vp(out))
    return out


class DeviceCoarse:
    """Device copy of a coarse space (``coarse.CoarseSpace``) and its coarse operators."""

    def __init__(self, ctx: Context, cs):
        self.ctx = ctx
        self.host = cs
        t = ctx.tensor
        self.node_agg = t(cs.node_agg)
        self.agg_ptr = t(cs.agg_ptr)
        self.agg_nodes = t(cs.agg_nodes)
        self.pair_ptr = t(cs.pair_ptr)
        self.pair_ij = t(np.ascontiguousarray(cs.pair_ij))
        self.pair_q = t(cs.pair_q)
        self.pair_row = t(cs.pair_row)
        self.P = t(np.ascontiguousarray(cs.P.reshape(-1)))
        self.P32 = _to_f32(ctx, self.P)
        xh = np.zeros((cs.P.shape[0], 3))
        xh[:, :cs.xhat.shape[1]] = cs.xhat
        self.xhat = t(np.ascontiguousarray(xh.reshape(-1)))
        self.xhat32 = _to_f32(ctx, self.xhat)
        self.fixed_bits = t(np.ascontiguousarray(cs.fixed_bits))
        self.T = t(np.ascontiguousarray(cs.T.reshape(-1)))
        self.T32 = _to_f32(ctx, self.T)
        self.n_coarse = cs.n_coarse
        self.ld = int(_hip.load().fe_coarse_ld(cs.n_coarse))
        # block-diagonal space of a batch: one dense coarse block per mesh
        if cs.agg_seg is not None:
            self.n_seg = cs.agg_seg.size - 1
            self.agg_seg = t(cs.agg_seg)
            self.seg_of_agg = t(cs.seg_of_agg if cs.seg_of_agg.size else np.zeros(1, np.int32))
            self.blk_ofs = t(cs.blk_ofs)
            self.blk_total = int(cs.blk_ofs[-1])
        else:
            self.n_seg = 0
            self.agg_seg = self.seg_of_agg = self.blk_ofs = None
            self.blk_total = 0

    def struct(self) -> _hip.FeCoarse:
        cs = self.host
        return _hip.FeCoarse(cs.P.shape[0], cs.n_agg, cs.n_pairs, cs.dof, cs.nc, _vp(self.node_agg),
                             _vp(self.agg_ptr), _vp(self.agg_nodes), _vp(self.pair_ptr), _vp(self.pair_ij),
                             _vp(self.pair_q), _vp(self.pair_row), _vp(self.P), _vp(self.P32),
                             _vp(self.xhat), _vp(self.xhat32), _vp(self.fixed_bits), _vp(self.T),
                             _vp(self.T32), self.n_seg, self.host.max_block, self.blk_total,
                             _vp(self.agg_seg), _vp(self.seg_of_agg), _vp(self.blk_ofs))

    def operator(self, A64: BsrMatrix, invert: bool = True, ctx: Optional[Context] = None):
        """(A_c or A_c^-1 [ld][ld] FP64, FP32 copy of the inverse or None) for the FP64 plane matrix A64,
        computed on ``ctx``'s stream (default: the coarse space's context)."""
        ctx = self.ctx if ctx is None else ctx
        # one system: [ld][ld]; a batch: the per-mesh blocks (blk_total values, no scratch)
        size = self.blk_total if self.n_seg else self.ld * self.ld
        Ac = torch.empty(max(size, 1), dtype=torch.float64, device=ctx.torch_device)
        A32 = torch.empty(max(size, 1), dtype=torch.float32, device=ctx.torch_device) if invert else None
        scratch = torch.empty(self.ld * self.ld + 2 * 64 * 64, dtype=torch.float64,
                              device=ctx.torch_device) if (invert and not self.n_seg) else None
        a = A64.struct()
        c = self.struct()
        _hip.call('fe_coarse_operator', ctx.h, ctypes.byref(a), ctypes.byref(c), int(invert), _vp(Ac),
                  _vp(A32), _vp(scratch))
        return Ac, A32

    def start_inverse(self, A64: BsrMatrix, ready: torch.cuda.Event):
        """Queue A_c^-1 (and its FP32 copy) for the FP64 plane matrix A64 on the side stream of A64's
        context, after the event ``ready``; returns (self, Ainv64, Ainv32, done event), or None when
        A64 is not an FP64 plane matrix.  The coarse operator reads constrained rows and columns as
        zero, so the assembled matrix gives the same A_c as its Dirichlet-eliminated copy: started
        from it, the inverse runs beside the elimination."""
        if A64.layout != _hip.FE_LAYOUT_PLANES or A64.precision != 0:
            return None
        side = A64.graph.ctx.side()
        side.stream.wait_event(ready)
        with torch.cuda.stream(side.stream):
            inv64, inv32 = self.operator(A64, invert=True, ctx=side)
        A64.vals.record_stream(side.stream)
        done = torch.cuda.Event()
        done.record(side.stream)
        return (self, inv64, inv32, done)

    def apply(self, Ainv: torch.Tensor, r: torch.Tensor, u: torch.Tensor):
        """u += P Ainv P^T r (FP64 or FP32 by the dtype of r; a block-diagonal space takes its FP64
        blocks as Ainv at either precision)."""
        prec = 0 if r.dtype == torch.float64 else 1
        if self.n_seg and Ainv.dtype != torch.float64:
            raise ValueError('a block-diagonal coarse space applies its FP64 blocks')
        el = 8 if prec == 0 else 4
        work = torch.empty(2 * ((self.n_coarse * el + 255) // 256) * 256, dtype=torch.uint8,
                           device=r.device)
        c = self.struct()
        _hip.call('fe_coarse_apply', self.ctx.h, ctypes.byref(c), prec, _vp(Ainv), _vp(r), _vp(u), _vp(work))
        return u


def _to_f32(ctx: Context, src: torch.Tensor) -> torch.Tensor:
    dst = torch.empty(src.numel(), dtype=torch.float32, device=src.device)
    cast_f64_to_f32(ctx, src, dst)
    return dst


class PcgSolver:
    """PCG on a Dirichlet-applied SPD matrix, with its workspace kept between solves.

    ``segments``: node-row offsets (n_seg + 1 entries) of independent blocks of a block-diagonal
    system (e.g. many plates assembled together); each block then iterates and stops on its own
    (``fe_pcg_batched``).  ``graph``: replay each chunk of iterations from a captured HIP graph.
    ``M64``: block preconditioner [n_nodes][d][d] FP64, a callable returning it, or None (the blocks
    of A64)."""

    def __init__(self, A64: BsrMatrix, precond: str = 'jacobi', mixed: bool = False,
                 M64: Optional[torch.Tensor] = None, A32: Optional[BsrMatrix] = None,
                 check_every: int = 16, ru_delta: float = 0.1, segments=None, graph: bool = False,
                 coarse: Optional['DeviceCoarse'] = None):
        self.ctx = A64.graph.ctx
        started = getattr(A64, 'coarse_inverse', None)  # queued by FESystem.prepare
        if A64.layout != _hip.FE_LAYOUT_PLANES or A64.precision != 0:
            A64 = A64.convert(_hip.FE_LAYOUT_PLANES, torch.float64)
        self.A64 = A64
        self.precond = precond
        self.mixed = bool(mixed)
        self.check_every = check_every
        self.ru_delta = ru_delta
        self.graph = bool(graph)
        self.coarse = coarse
        self.Ainv64 = self.Ainv32 = None
        if coarse is not None:
            nseg = 1 if segments is None else len(segments) - 1
            if (segments is None) != (coarse.n_seg == 0) or (coarse.n_seg and coarse.n_seg != nseg):
                raise ValueError('a batched solve needs the block-diagonal coarse space of its meshes '
                                 '(one block per segment), a single system the plain one')
            # the coarse operator and its inverse (part of the preconditioner set-up of every
            # solver) run on the side stream, beside the block preconditioner and the iteration
            # matrix built below on the main stream -- or were queued there already, from the
            # assembled matrix, by FESystem.prepare
            if started is not None and started[0] is coarse:
                _, self.Ainv64, self.Ainv32, coarse_done = started
            else:
                side = self.ctx.side()
                ready = torch.cuda.Event()
                ready.record(self.ctx.stream)
                side.stream.wait_event(ready)
                with torch.cuda.stream(side.stream):
                    self.Ainv64, self.Ainv32 = coarse.operator(A64, invert=True, ctx=side)
                coarse_done = torch.cuda.Event()
                coarse_done.record(side.stream)
        # block preconditioner: given, computed by a callable (e.g. a learned model's inference), or
        # the blocks of A64
        if M64 is None:
            M64 = precond_setup(A64, 'jacobi' if precond == 'jacobi' else 'block_jacobi')
        elif callable(M64):
            M64 = M64()
        self.M64 = M64
        # the iteration matrix: a copy on the padded pattern (16-byte loads in the SpMV), FP32 with
        # mixed precision; A64 keeps the true residual
        if self.mixed:
            self.A32 = A64.padded(torch.float32) if A32 is None else A32
            self.M32 = _to_f32(self.ctx, self.M64)
        else:
            self.A32 = A64.padded(torch.float64) if A32 is None else A32
            self.M32 = None
        if coarse is not None:
            self.ctx.stream.wait_event(coarse_done)
            for t in (self.Ainv64, self.Ainv32):
                t.record_stream(self.ctx.stream)
        if segments is not None:
            seg = np.ascontiguousarray(np.asarray(segments, dtype=np.int32))
            if seg.ndim != 1 or seg.size < 2 or seg[0] != 0 or seg[-1] != A64.graph.n_nodes:
                raise ValueError('segments must be node offsets from 0 to n_nodes')
            self.segments = seg
        else:
            self.segments = None
        self._work = None

    def set_preconditioner(self, M64: torch.Tensor):
        self.M64 = M64
        if self.mixed:
            self.M32 = _to_f32(self.ctx, M64)

    def _opts(self, tol, max_iter, profile=False):
        pc = _hip.FE_PC_JACOBI if self.precond == 'jacobi' else _hip.FE_PC_BLOCK_JACOBI
        return _hip.FePcgOpts(float(tol), int(max_iter), pc, int(self.mixed), int(self.check_every),
                              float(self.ru_delta), int(profile), int(self.graph))

    def solve(self, b: torch.Tensor, x: Optional[torch.Tensor] = None, tol: float = 1e-8,
              max_iter: Optional[int] = None, profile: bool = False):
        """Solve from initial guess x (zeros if None); returns (x, info dict).  With segments, info
        also holds per-segment arrays (iterations, status, relres, ...)."""
        ctx = self.ctx
        if max_iter is None:
            max_iter = max(2000, min(10 * self.A64.n, 50000))
        if x is None:
            x = torch.zeros_like(b)
        opts = self._opts(tol, max_iter, profile)
        a64 = self.A64.struct()
        a32 = self.A32.struct() if self.A32 is not None else None
        nbytes = ctypes.c_int64()
        cstruct = self.coarse.struct() if self.coarse is not None else None
        if self.segments is not None:
            nseg = self.segments.size - 1
            segp = self.segments.ctypes.data_as(ctypes.c_void_p)
        if cstruct is not None and self.segments is None:
            _hip.call('fe_pcg_coarse_workspace', ctx.h, ctypes.byref(a64), ctypes.byref(cstruct),
                      ctypes.byref(opts), ctypes.byref(nbytes))
        elif cstruct is not None:
            _hip.call('fe_pcg_batched_coarse_workspace', ctx.h, ctypes.byref(a64), nseg, segp,
                      ctypes.byref(cstruct), ctypes.byref(opts), ctypes.byref(nbytes))
        elif self.segments is None:
            _hip.call('fe_pcg_workspace', ctx.h, ctypes.byref(a64), ctypes.byref(opts),
                      ctypes.byref(nbytes))
        else:
            _hip.call('fe_pcg_batched_workspace', ctx.h, ctypes.byref(a64), nseg, segp,
                      ctypes.byref(opts), ctypes.byref(nbytes))
        if self._work is None or self._work.numel() < nbytes.value:
            self._work = torch.empty(nbytes.value, dtype=torch.uint8, device=ctx.torch_device)
        stats = _hip.FePcgStats()
        a32p = ctypes.byref(a32) if a32 is not None else None
        seg_info = None
        if cstruct is not None and self.segments is None:
            _hip.call('fe_pcg_coarse', ctx.h, ctypes.byref(a64), a32p, _vp(self.M64), _vp(self.M32),
                      ctypes.byref(cstruct), _vp(self.Ainv64), _vp(self.Ainv32), _vp(b), _vp(x),
                      _vp(self._work), self._work.numel(), ctypes.byref(opts), ctypes.byref(stats))
        elif self.segments is None:
            _hip.call('fe_pcg', ctx.h, ctypes.byref(a64), a32p, _vp(self.M64), _vp(self.M32), _vp(b),
                      _vp(x), _vp(self._work), self._work.numel(), ctypes.byref(opts),
                      ctypes.byref(stats))
        else:
            seg_stats = (_hip.FePcgSegStats * nseg)()
            if cstruct is not None:
                _hip.call('fe_pcg_batched_coarse', ctx.h, ctypes.byref(a64), a32p, _vp(self.M64),
                          _vp(self.M32), nseg, segp, ctypes.byref(cstruct), _vp(self.Ainv64),
                          _vp(self.Ainv32), _vp(b), _vp(x), _vp(self._work), self._work.numel(),
                          ctypes.byref(opts), ctypes.byref(stats), ctypes.cast(seg_stats, ctypes.c_void_p))
            else:
                _hip.call('fe_pcg_batched', ctx.h, ctypes.byref(a64), a32p, _vp(self.M64), _vp(self.M32),
                          nseg, segp, _vp(b), _vp(x), _vp(self._work), self._work.numel(),
                          ctypes.byref(opts), ctypes.byref(stats), ctypes.cast(seg_stats, ctypes.c_void_p))
            seg_info = {f: np.array([getattr(q, f) for q in seg_stats])
                        for f in ('iterations', 'reliable_updates', 'status', 'restarts', 'fallback',
                                  'relres', 'relres_rec', 'bnorm')}
        info = {'iterations': stats.iterations, 'reliable_updates': stats.reliable_updates,
                'restarts': stats.restarts,
                'status': stats.status, 'relres': stats.relres, 'relres_rec': stats.relres_rec,
                'bnorm': stats.bnorm, 'spmv_ms': stats.spmv_ms, 'spmv_samples': stats.spmv_samples,
                'fallbacks': stats.fallbacks}
        if seg_info is not None:
            info['segments'] = seg_info
        return x, info
