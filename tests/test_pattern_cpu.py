"""Host pattern builder (fe_pattern_create, runs on the CPU) against scipy's COO->CSR -- the
reference's own call (finite_elements/analysis.py:661) -- bit-exact, and CPU emulations of the two
assembly maps (sorted contribution map and block-level map) against scipy's summed values."""

import os

import numpy as np
import pytest

from finite_elements_amd import meshgen
from finite_elements_amd.device import Pattern, element_schedule
from finite_elements_amd.gmsh import read_msh
from oracle import vectorized as vec
from tests.conftest import GOLDEN, rowscaled_err


def _cases():
    rng = np.random.default_rng(1)
    out = []
    c, t = meshgen.structured_tri(40, 10, 1.0, 0.25)
    out.append(('cantilever', 'tri_elastic', c, t))
    c, t = meshgen.structured_tri(17, 13)
    out.append(('magnetic', 'tri_magnetic', c, t))
    c, t = meshgen.kuhn_tets(4)
    out.append(('kuhn4', 'tet_elastic', c, t))
    # a permuted node numbering (caller-supplied nodes / node_to_index are honoured verbatim)
    perm = rng.permutation(c.shape[0])
    inv = np.argsort(perm)
    out.append(('kuhn4_perm', 'tet_elastic', c[perm], inv[t].astype(np.int32)))
    g = np.load(os.path.join(GOLDEN, 'gmsh_meshes.npz'))
    out.append(('beam_2d_0.3', 'tri_elastic', g['beam_2d_0.3_coords'][:, :2], g['beam_2d_0.3_triangles']))
    out.append(('cube', 'tet_elastic', g['cube_coords'], g['cube_tetrahedra']))
    return out


def _materials(kind, ne, rng):
    if kind == 'tri_magnetic':
        return {'mu': rng.uniform(1e-6, 1e-3, ne)}
    m = {'E': rng.uniform(1e9, 3e11, ne), 'nu': rng.uniform(0.2, 0.4, ne), 'rho': rng.uniform(1e3, 8e3, ne)}
    if kind == 'tri_elastic':
        m['thickness'] = rng.uniform(0.01, 1.0, ne)
    return m


DOF = {'tri_elastic': 2, 'tet_elastic': 3, 'tri_magnetic': 1}


@pytest.mark.parametrize('name,kind,coords,conn', _cases(), ids=[c[0] for c in _cases()])
def test_pattern_bit_exact_and_maps(name, kind, coords, conn):
    rng = np.random.default_rng(7)
    d = DOF[kind]
    N = coords.shape[0]
    mats = _materials(kind, conn.shape[0], rng)
    plane = 2 if kind == 'tri_elastic' else 0
    K = vec.assemble_csr(kind, coords, conn, mats, plane)
    pat = Pattern(conn, d, N, sym=True, with_gather=True)
    # pattern: bit-exact with scipy (sorted columns, explicit zeros kept)
    assert pat.nnz == K.nnz
    assert np.array_equal(pat.rowptr, K.indptr.astype(np.int32))
    assert np.array_equal(pat.colind, K.indices.astype(np.int32))
    # two-phase map: sum the symmetric-storage element matrices through gather/seg_ptr
    ke = vec.element_matrices(kind, coords, conn, mats, plane)  # (E, m*m)
    m = int(round(np.sqrt(ke.shape[1])))
    iu = np.triu_indices(m)
    ke_sym = ke.reshape(-1, m, m)[:, iu[0], iu[1]].T.copy()  # entry-major [msym][E]
    flat = ke_sym.reshape(-1)
    contrib = flat[pat.gather]
    seg = np.repeat(np.arange(pat.nnz), np.diff(pat.seg_ptr))
    vals = np.zeros(pat.nnz)
    np.add.at(vals, seg, contrib)
    got = K.copy()
    got.data = vals
    assert rowscaled_err(K, got) < 1e-13
    # segments list contributions in ascending element order
    e_of = pat.gather % conn.shape[0]
    same = seg[1:] == seg[:-1]
    assert np.all(e_of[1:][same] >= e_of[:-1][same])
    # block-level map of the fused assembly
    bseg_owner = np.repeat(np.arange(pat.nnzb), np.diff(pat.bseg_ptr))
    node_of_block = np.repeat(np.arange(N), np.diff(pat.bptr))
    i = node_of_block[bseg_owner]
    kl = pat.bseg.astype(np.int64) >> 2
    b = pat.bseg.astype(np.int64) & 3
    packed = pat.inc[pat.inc_ptr[i] + kl]
    e = packed // conn.shape[1]
    a = packed % conn.shape[1]
    assert np.all(conn[e, a] == i)
    assert np.all(conn[e, b] == pat.bcol[bseg_owner])
    K3 = ke.reshape(-1, conn.shape[1], d, conn.shape[1], d)
    bsr = np.zeros((pat.nnzb, d, d))
    np.add.at(bsr, bseg_owner, K3[e, a, :, b, :])
    # BSR -> CSR order inside each node chunk
    csr_vals = np.empty(pat.nnz)
    for node in range(N):
        b0, b1 = pat.bptr[node], pat.bptr[node + 1]
        chunk = bsr[b0:b1]  # (deg, d, d)
        csr_vals[b0 * d * d:b1 * d * d] = chunk.transpose(1, 0, 2).reshape(-1)
    got.data = csr_vals
    assert rowscaled_err(K, got) < 1e-13


@pytest.mark.parametrize('name,kind,coords,conn', _cases(), ids=[c[0] for c in _cases()])
@pytest.mark.parametrize('max_inc', [64, 256])
def test_element_schedule(name, kind, coords, conn, max_inc):
    """CPU emulation of the fused assembly's element schedule: every block contribution resolves,
    through its incidence's packed slot code, to the element and local vertex the incidence names,
    and each workgroup's element list is the ascending set of its incidences' elements."""
    d = DOF[kind]
    N = coords.shape[0]
    npe = conn.shape[1]
    pat = Pattern(conn, d, N)
    cap = max(max_inc, int(np.diff(pat.inc_ptr).max()))
    blk, biggest = Pattern.schedule_for(pat.inc_ptr, cap)
    wptr, welem, inc_sa, max_elem = element_schedule(pat, blk)
    assert wptr[0] == 0 and wptr[-1] == welem.size
    assert max_elem == np.diff(wptr).max()
    for w in range(blk.size - 1):
        i0, i1 = pat.inc_ptr[blk[w]], pat.inc_ptr[blk[w + 1]]
        elems = pat.inc[i0:i1] // npe
        mine = welem[wptr[w]:wptr[w + 1]]
        assert np.array_equal(mine, np.unique(elems))
        slot = inc_sa[i0:i1].astype(np.int64) >> 2
        a = inc_sa[i0:i1].astype(np.int64) & 3
        assert np.array_equal(mine[slot], elems)
        assert np.array_equal(a, pat.inc[i0:i1] % npe)
    # phase 2 emulation: block (i, j) <- element (slot, a, b) via bseg and inc_sa
    owner = np.repeat(np.arange(pat.nnzb), np.diff(pat.bseg_ptr))
    i = np.repeat(np.arange(N), np.diff(pat.bptr))[owner]
    t = pat.inc_ptr[i] + (pat.bseg.astype(np.int64) >> 2)
    w_of_node = np.repeat(np.arange(blk.size - 1), np.diff(blk))
    e = welem[wptr[w_of_node[i]] + (inc_sa[t].astype(np.int64) >> 2)]
    a = inc_sa[t].astype(np.int64) & 3
    b = pat.bseg.astype(np.int64) & 3
    assert np.all(conn[e, a] == i) and np.all(conn[e, b] == pat.bcol[owner])


@pytest.mark.parametrize('name,kind,coords,conn', _cases(), ids=[c[0] for c in _cases()])
@pytest.mark.parametrize('max_inc', [64, 256])
def test_node_lists_of_the_schedule(name, kind, coords, conn, max_inc):
    """fe_schedule_nodes (the wave-specialised assembly's node lists): each workgroup's list is the
    ascending set of its elements' nodes, every element vertex resolves through its local position
    to its own node, positions past the element's vertex count are 0, and the header words 10 and
    11 hold each list's slice."""
    from finite_elements_amd.device import Context, DeviceGraph
    pat = Pattern(conn, DOF[kind], coords.shape[0])
    cap = max(max_inc, int(np.diff(pat.inc_ptr).max()))
    blk, biggest = Pattern.schedule_for(pat.inc_ptr, cap)
    g = DeviceGraph.__new__(DeviceGraph)  # host half only: the arrays are built before any upload
    g.pattern, g.n_nodes = pat, pat.n_nodes
    g.ctx = type('HostCtx', (), {'tensor': staticmethod(lambda a: a)})()
    g._conn = np.ascontiguousarray(conn.astype(np.int32))
    g.set_schedule(blk, biggest, 1024, 8)
    npe = conn.shape[1]
    hdr = g.wg_hdr.reshape(-1, 12)
    wptr = g._welem_ptr_host
    local = g.welem_local.view(np.uint16).reshape(-1, 4)
    assert g.max_wnode == int(hdr[:, 11].max())
    for w in range(blk.size - 1):
        n0, nn = hdr[w, 10], hdr[w, 11]
        nodes = g.wnode[n0:n0 + nn]
        elems = g._welem_host[wptr[w]:wptr[w + 1]]
        assert np.array_equal(nodes, np.unique(conn[elems]))
        loc = local[wptr[w]:wptr[w + 1]].astype(np.int64)
        assert np.array_equal(nodes[loc[:, :npe]], conn[elems])
        assert np.all(loc[:, npe:] == 0)
    # fe_schedule_codes: every contribution's resolved code names the element and the two local
    # vertices that the block's row and column nodes are, and brow is the row's place in its range
    code = g.wcode.view(np.uint16).astype(np.int64)
    brow = g.brow.view(np.uint16).astype(np.int64)
    N = pat.n_nodes
    rows = np.repeat(np.arange(N), np.diff(pat.bptr))
    w_of_node = np.repeat(np.arange(blk.size - 1), np.diff(blk))
    assert np.array_equal(brow & 0x7fff, rows - blk[w_of_node[rows]])
    assert np.array_equal(brow >> 15, (pat.bcol == rows).astype(np.int64))
    owner = np.repeat(np.arange(pat.nnzb), np.diff(pat.bseg_ptr))
    i = rows[owner]
    e = g._welem_host[wptr[w_of_node[i]] + (code >> 4)]
    a, b = (code >> 2) & 3, code & 3
    assert np.all(conn[e, a] == i) and np.all(conn[e, b] == pat.bcol[owner])


def test_schedule_covers_nodes():
    c, t = meshgen.kuhn_tets(6)
    pat = Pattern(t, 3, c.shape[0])
    assert pat.blk[0] == 0 and pat.blk[-1] == c.shape[0]
    assert np.all(np.diff(pat.blk) > 0) and np.all(np.diff(pat.blk) <= 256)
    inc = pat.inc_ptr[pat.blk[1:]] - pat.inc_ptr[pat.blk[:-1]]
    assert inc.max() == pat.max_inc <= 192
    blk2, biggest = Pattern.schedule_for(pat.inc_ptr, 192)
    assert np.array_equal(blk2, pat.blk) and biggest == pat.max_inc
    with pytest.raises(ValueError):
        Pattern.schedule_for(pat.inc_ptr, 5)
    # fe_schedule_ranges against the greedy rule written out: a range takes consecutive nodes while
    # it stays within max_inc incidences and max_nodes nodes
    g = np.load(os.path.join(GOLDEN, 'gmsh_meshes.npz'))
    for conn, n in ((t, c.shape[0]), (g['cube_tetrahedra'], g['cube_coords'].shape[0])):
        p = Pattern(conn, 3, n)
        counts = np.diff(p.inc_ptr).tolist()
        for max_inc, max_nodes in ((64, 256), (96, 7), (300, 256)):
            want, cur, nodes, big = [0], 0, 0, 0
            for i, k in enumerate(counts):
                if nodes and (cur + k > max_inc or nodes >= max_nodes):
                    want.append(i)
                    big = max(big, cur)
                    cur = nodes = 0
                cur += k
                nodes += 1
            want.append(len(counts))
            big = max(big, cur)
            got, got_big = Pattern.schedule_for(p.inc_ptr, max_inc, max_nodes)
            assert np.array_equal(got, want) and got_big == big


@pytest.mark.parametrize('case,expect', [
    ('cfg1', (902, 11804)), ('cfg2', (203401, 1420201)), ('cfg5', (352947, 15369273)),
    ('cfg4', (823875, 36168777))])
def test_config_sizes(case, expect):
    if case == 'cfg1':
        c, t = meshgen.structured_tri(40, 10, 1.0, 0.25)
        d = 2
    elif case == 'cfg2':
        c, t = meshgen.structured_tri(450, 450, 2.0, 2.0, -1.0, -1.0)
        d = 1
    elif case == 'cfg5':
        c, t = meshgen.kuhn_tets(48)
        d = 3
    else:
        c, t = meshgen.kuhn_tets(64)
        d = 3
    pat = Pattern(t, d, c.shape[0])
    assert (c.shape[0] * d, pat.nnz) == expect


def test_gmsh_fixture_counts():
    g = np.load(os.path.join(GOLDEN, 'gmsh_meshes.npz'))
    counts = {k: g[k].shape[0] for k in g.files}
    assert counts['beam_2d_0.8_coords'] == 43 and counts['beam_2d_0.8_triangles'] == 54
    assert counts['beam_2d_0.1_coords'] == 1302 and counts['beam_2d_0.1_triangles'] == 2382
    assert counts['beam3d_3_tetrahedra'] == 5199 and counts['cube_tetrahedra'] == 2576
