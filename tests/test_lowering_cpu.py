"""Host-side lowering of loads and Dirichlet conditions (FiniteElementAnalysis, no GPU needed)
against the oracle's restatement of finite_elements/analysis.py:28-110, 201-279, 407-543, 665-739,
on hand-built cases that exercise the reference's quirks: last write wins per (node, dimension),
first-insertion order, element loads split by area with factors (0, A, A), element conditions scaled
by those factors and edge conditions / loads by 0.5."""

import numpy as np
import pytest

import finite_elements_amd as fe
from finite_elements_amd import conditions, elements, loads, lowering, mesh, meshgen
from finite_elements_amd.analysis import FiniteElementAnalysis
from oracle import reference as ref

MU0 = 4e-7 * np.pi


class Edge:
    def __init__(self, start, end):
        self.start, self.end = start, end


def magnetic_beam():
    """scripts/Magnetic/finite_element_beam.py:14-66."""
    mus = [MU0 * 100000, MU0, MU0 * 50000]
    groups, o_tris, o_mats = [], [], []
    for phase in range(3):
        els = []
        for i in range(3 * phase, 3 * phase + 3):
            t1 = mesh.TriangularElement2D([mesh.Point2D(i, 0), mesh.Point2D(i + 1, 0), mesh.Point2D(i, 1)])
            t2 = mesh.TriangularElement2D([mesh.Point2D(i + 1, 1), mesh.Point2D(i + 1, 0), mesh.Point2D(i, 1)])
            els += [fe.elements.MagneticElement2D(t1, mus[phase]), fe.elements.MagneticElement2D(t2, mus[phase])]
            o_tris += [tuple(p._c for p in t1.points), tuple(p._c for p in t2.points)]
            o_mats += [{'mu': mus[phase]}] * 2
        groups.append(mesh.ElementsGroup(els, f'phase{phase}'))
    m = mesh.Mesh(groups)
    return m, groups, o_tris, o_mats


def test_mesh_node_order_and_dof_map():
    m, _, _, _ = magnetic_beam()
    # first-appearance order over groups / elements / points
    assert m.nodes[0] == mesh.Point2D(0, 0) and m.nodes[1] == mesh.Point2D(1, 0)
    assert m.node_to_index[mesh.Point2D(0, 1)] == 2
    assert fe.core.global_matrix_positions(2, 3) == ref.global_matrix_positions(2, 3)


def test_magnetic_element_load_lowering():
    m, groups, o_tris, o_mats = magnetic_beam()
    el_loads = [loads.ElementsLoad(groups[0].elements[0:2], 1e10, 1)]
    bcs = [conditions.NodeBoundaryCondition(n, 0, 1) for n in m.nodes if abs(n[0] - 9) < 1e-6]
    an = FiniteElementAnalysis(m, el_loads, [], [], [], [], bcs, [], [])
    F = an.create_source_matrix()
    prob = ref.Problem('tri_magnetic', [n._c for n in m.nodes], list(zip(o_tris, o_mats)),
                       element_loads=[([0, 1], 1e10, 1)],
                       node_bcs=[(n._c, 0.0, 1) for n in m.nodes if abs(n[0] - 9) < 1e-6])
    Fo = ref.create_source_matrix(prob)
    assert F.shape == Fo.shape
    np.testing.assert_allclose(F, Fo, rtol=0, atol=1e-9 * np.abs(Fo).max())
    # factors (0, A, A) with last write wins: element 1 = [(1,1), (1,0), (0,1)] is written last, so
    # (1,0) and (0,1) carry A-weighted values, (1,1) gets ~0 and (0,0) keeps element 0's ~0
    assert np.count_nonzero(np.abs(F[:, 0]) > 1e-3 * np.abs(F).max()) == 2


def _two_groups_tri():
    pts = [mesh.Node2D(x, y) for x, y in [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1)]]
    P = lambda i: pts[i]  # noqa: E731
    tri = [(0, 1, 4), (0, 4, 3), (1, 2, 5), (1, 5, 4)]
    els = [fe.elements.ElasticityTriangularElement2D(mesh.TriangularElement2D([P(a), P(b), P(c)]),
                                                     210e9, 0.3, 7850, 0.01) for a, b, c in tri]
    m = mesh.Mesh([mesh.ElementsGroup(els[:2], 'a'), mesh.ElementsGroup(els[2:], 'b')])
    o_els = [(tuple(P(i)._c for i in t), {'E': 210e9, 'nu': 0.3, 'thickness': 0.01, 'rho': 7850})
             for t in tri]
    return m, pts, els, o_els


def test_elastic_lowering_quirks():
    m, pts, els, o_els = _two_groups_tri()
    # duplicated node loads: the second value wins; edge load halves onto its endpoints and
    # overwrites the node (4, dim 2) set before it
    nl = [loads.NodeLoad(pts[5], -10.0, 2), loads.NodeLoad(pts[5], -25.0, 2), loads.NodeLoad(pts[4], 3.0, 1),
          loads.NodeLoad(pts[4], 7.0, 2)]
    edl = [loads.EdgeLoad(Edge(pts[4], pts[5]), 8.0, 2)]
    el_l = [loads.ElementsLoad([els[2], els[3]], 100.0, 1)]
    nbc = [conditions.NodeBoundaryCondition(pts[0], 0.0, 1), conditions.NodeBoundaryCondition(pts[0], 0.5, 1),
           conditions.NodeBoundaryCondition(pts[3], 0.0, 2)]
    ebc = [conditions.ElementBoundaryCondition(els[1], 2.0, 1)]
    edbc = [conditions.EdgeBoundaryCondition(Edge(pts[0], pts[3]), 4.0, 2)]
    an = FiniteElementAnalysis(m, el_l, edl, nl, [], [], nbc, edbc, ebc, plane_strain=False,
                               plane_stress=True)
    prob = ref.Problem('tri_elastic', [n._c for n in m.nodes], o_els,
                       node_loads=[(pts[5]._c, -10.0, 2), (pts[5]._c, -25.0, 2), (pts[4]._c, 3.0, 1),
                                   (pts[4]._c, 7.0, 2)],
                       edge_loads=[((pts[4]._c, pts[5]._c), 8.0, 2)],
                       element_loads=[([2, 3], 100.0, 1)],
                       node_bcs=[(pts[0]._c, 0.0, 1), (pts[0]._c, 0.5, 1), (pts[3]._c, 0.0, 2)],
                       element_bcs=[(1, 2.0, 1)],
                       edge_bcs=[((pts[0]._c, pts[3]._c), 4.0, 2)],
                       plane_strain=False, plane_stress=True)
    F = an.create_source_matrix()
    Fo = ref.create_source_matrix(prob)
    assert F.shape == Fo.shape
    np.testing.assert_allclose(F, Fo, rtol=0, atol=1e-12 * np.abs(Fo).max())
    d, r, c = an.c_matrix_boundary_conditions()
    positions = ref.global_matrix_positions(2, len(prob.nodes))
    do, ro, co = ref.c_matrix_boundary_conditions(prob, positions, prob.node_to_index())
    assert (d, r, c) == (do, ro, co)
    assert an.get_source_matrix_length() == len(Fo)
    # row/col indices of one element
    assert an.get_row_col_indices(els[3]) == ref.get_row_col_indices(
        prob, o_els[3], positions, prob.node_to_index())


def test_continuity_conditions_need_a_scalar_problem():
    """The reference writes node indices as DOF rows (analysis.py:190-197), which only means
    something for one DOF per node: vector problems are refused."""
    m, pts, els, _ = _two_groups_tri()
    an = FiniteElementAnalysis(m, [], [], [], [], [conditions.ContinuityCondition(pts[0], pts[1], 1)],
                               [], [], [], plane_strain=False, plane_stress=True)
    with pytest.raises(NotImplementedError):
        an.c_matrix_continuity_conditions()


def _magnet_mesh(n=8):
    """Magnetostatic triangles on [0,1]^2 with a magnet region [0.25, 0.75]^2 (mesh objects)."""
    coords, conn = meshgen.structured_tri(n, n)
    am = mesh.ArrayMesh(coords, conn, 'tri_magnetic', {'mu': MU0})
    els = am.elements_groups[0].elements
    cen = coords[conn].mean(axis=1)
    inside = np.nonzero((np.abs(cen[:, 0] - 0.5) < 0.25) & (np.abs(cen[:, 1] - 0.5) < 0.25))[0]
    return am, coords, conn, els, inside


def test_magnet_sources_closed_form():
    """A uniformly magnetised square: the sources are the equivalent surface current K = M x n_out
    lumped on the contour nodes: + My h on the left side, - My h on the right, - Mx h at the bottom,
    + Mx h at the top (corners: half of each side), zero inside, zero in total."""
    n, h = 8, 1.0 / 8
    am, coords, conn, els, inside = _magnet_mesh(n)
    M = (0.7, -1.3)
    verts = [tuple(tuple(coords[i]) for i in conn[e]) for e in inside]
    src = ref.magnet_sources(verts, [], M)
    f = {}
    for p, v in src:
        f[p] = f.get(p, 0.0) + v
    want = {}
    lo, hi = 0.25, 0.75
    for (x, y) in {p for p, _ in src}:
        w = 0.0
        on_l, on_r, on_b, on_t = np.isclose(x, lo), np.isclose(x, hi), np.isclose(y, lo), np.isclose(y, hi)
        side_y = 0.5 if (on_b or on_t) else 1.0
        side_x = 0.5 if (on_l or on_r) else 1.0
        if on_l:
            w += M[1] * h * side_y
        if on_r:
            w -= M[1] * h * side_y
        if on_b:
            w -= M[0] * h * side_x
        if on_t:
            w += M[0] * h * side_x
        want[(x, y)] = w
    assert set(f) == set(want)
    for p in f:
        assert abs(f[p] - want[p]) <= 1e-14
    assert abs(sum(f.values())) <= 1e-13
    # the nodes strictly inside the magnet receive nothing
    assert all(lo < p[0] < hi and lo < p[1] < hi for p in set(map(tuple, coords[conn[inside]].reshape(-1, 2)))
               if p not in f)


def test_magnet_load_lowering_matches_oracle():
    """FiniteElementAnalysis with MagnetLoads (one with non-contour nodes) and an ElementsLoad: the
    source matrix equals the oracle's per-edge restatement, and the magnet rows come in its order."""
    am, coords, conn, els, inside = _magnet_mesh(10)
    nodes = am.nodes
    mag_els = [els[e] for e in inside]
    # the magnet's left side declared non-contour: those edges drop out
    left = [nodes[i] for i in range(coords.shape[0]) if np.isclose(coords[i, 0], 0.3)]
    ml1 = loads.MagnetLoad(mag_els, [], (0.2, 1.1))
    ml2 = loads.MagnetLoad(mag_els[:7], left, (-0.5, 0.25))
    coil = loads.ElementsLoad([els[e] for e in range(5)], 3.0, 1)
    bcs = [conditions.NodeBoundaryCondition(nodes[i], 0.0, 1) for i in range(11)]
    an = FiniteElementAnalysis(am, [coil], [], [], [ml1, ml2], [], bcs, [], [])
    data, row = an.source_c_matrix_magnet_loads()
    want = []
    for ml in (ml1, ml2):
        verts = [tuple(tuple(p._c) for p in el.points) for el in ml.elements]
        inner = [tuple(p._c) for p in ml.non_contour_nodes]
        want += ref.magnet_sources(verts, inner, ml.magnetization_vector)
    idx = {tuple(coords[i]): i for i in range(coords.shape[0])}
    assert row == [idx[p] for p, _ in want]
    np.testing.assert_allclose(data, [v for _, v in want], rtol=1e-15, atol=0)
    F = an.create_source_matrix()
    f = np.zeros(an.n_dofs)
    prob = ref.problem_from_arrays('tri_magnetic', coords, conn, {'mu': MU0})
    prob.element_loads = [(list(range(5)), 3.0, 1)]
    for p, v, _ in ref.collapsed_node_loads(prob):
        f[idx[p]] += v
    for p, v in want:
        f[idx[p]] += v
    np.testing.assert_allclose(F[:an.n_dofs, 0], f, rtol=0, atol=1e-15 * np.abs(f).max())


def test_vectorised_lowering_matches_objects():
    """The array lowering (analysis.load_arrays / boundary_condition_dofs) against the per-object
    reference semantics (the oracle) on a cantilever with every kind of load and condition, repeated
    keys included: exact equality of rows and values."""
    pr = meshgen.cantilever_2d(12, 4)
    am = mesh.ArrayMesh(pr.coords, pr.conn, 'tri_elastic', pr.materials)
    nodes, els = am.nodes, am.elements_groups[0].elements
    rng = np.random.default_rng(5)
    nl = [loads.NodeLoad(nodes[int(i)], float(v), int(d)) for i, v, d in
          zip(rng.integers(0, len(nodes), 40), rng.standard_normal(40), rng.integers(1, 3, 40))]
    el_loads = [loads.ElementsLoad([els[int(e)] for e in rng.choice(len(els), 9, replace=False)],
                                   float(rng.standard_normal()), int(d)) for d in (1, 2, 1)]
    edges = [loads.EdgeLoad(Edge(nodes[int(a)], nodes[int(b)]), float(v), int(d)) for a, b, v, d in
             zip(rng.integers(0, len(nodes), 6), rng.integers(0, len(nodes), 6), rng.standard_normal(6),
                 rng.integers(1, 3, 6))]
    nbc = [conditions.NodeBoundaryCondition(nodes[int(i)], float(v), int(d)) for i, v, d in
           zip(rng.integers(0, 13, 20), rng.standard_normal(20), rng.integers(1, 3, 20))]
    ebc = [conditions.ElementBoundaryCondition(els[int(e)], float(v), 2) for e, v in
           zip(rng.integers(0, len(els), 4), rng.standard_normal(4))]
    gbc = [conditions.EdgeBoundaryCondition(Edge(nodes[0], nodes[13]), 0.25, 1)]
    an = FiniteElementAnalysis(am, el_loads, edges, nl, [], [], nbc, gbc, ebc, plane_strain=False,
                               plane_stress=True)
    # object path: NodeLoad / NodeBoundaryCondition lists collapsed by the dictionaries
    data, row = an.source_c_matrix_loads()
    rows, vals = an.load_arrays()
    assert rows.tolist() == row and vals.tolist() == data
    bcs = an._collapsed_boundary_conditions()
    pos, val = an.boundary_condition_dofs()
    assert pos.tolist() == [an._dof(b.application, b.dimension) for b in bcs]
    assert val.tolist() == [b.source_c_matrix() for b in bcs]


def test_lowering_450_magnetostatics_is_fast():
    """The 450^2 coils through FiniteElementAnalysis (ElementsLoad on both coils, Dirichlet on the
    1,800 boundary nodes): lowered in well under a second, equal to the bench's array lowering."""
    import time
    pr = meshgen.magnetostatic_2d(n=450)
    am = mesh.ArrayMesh(pr.coords, pr.conn, 'tri_magnetic', pr.materials)
    cen = pr.coords[pr.conn].mean(axis=1)
    coil_p = np.nonzero((cen[:, 0] >= -0.25) & (cen[:, 0] <= -0.15) & (np.abs(cen[:, 1]) <= 0.1))[0]
    coil_n = np.nonzero((cen[:, 0] >= 0.15) & (cen[:, 0] <= 0.25) & (np.abs(cen[:, 1]) <= 0.1))[0]
    # element objects only for the coil elements (an ArrayMesh builds its own list on demand)
    sub_p = [elements.MagneticElement2D(mesh.TriangularElement2D([am.nodes[i] for i in pr.conn[e]]),
                                        float(pr.materials['mu'][e])) for e in coil_p]
    sub_n = [elements.MagneticElement2D(mesh.TriangularElement2D([am.nodes[i] for i in pr.conn[e]]),
                                        float(pr.materials['mu'][e])) for e in coil_n]
    for el, e in zip(sub_p + sub_n, np.concatenate([coil_p, coil_n])):
        el.mesh_index = int(e)
    bcs = [conditions.NodeBoundaryCondition(am.nodes[i], 0.0, 1) for i in pr.fixed_dofs]
    t0 = time.perf_counter()
    an = FiniteElementAnalysis(am, [loads.ElementsLoad(sub_p, 1e6, 1), loads.ElementsLoad(sub_n, -1e6, 1)],
                               [], [], [], [], bcs, [], [])
    f = an.load_vector()
    pos, val = an.boundary_condition_dofs()
    dt = time.perf_counter() - t0
    assert dt < 1.0, dt
    # the bench's problem arrays use the exact factors (0, A, A); the reference's formula rounds
    # them (a vertex-0 factor comes out near 1e-17 A, not 0)
    np.testing.assert_allclose(f, pr.f, rtol=0, atol=1e-12 * np.abs(pr.f).max())
    assert pos.tolist() == pr.fixed_dofs.tolist() and not val.any()


def test_plane_flag_errors():
    m, _, els, _ = _two_groups_tri()
    an = FiniteElementAnalysis(m, [], [], [], [], [], [], [], [], plane_strain=True, plane_stress=True)
    with pytest.raises(ValueError, match='just one'):
        an._plane_mode()
    an = FiniteElementAnalysis(m, [], [], [], [], [], [], [], [])
    with pytest.raises(ValueError, match='must be True'):
        an._plane_mode()


def test_last_write_wins_vectorised_matches_dict():
    rng = np.random.default_rng(3)
    nodes = rng.integers(0, 20, 200)
    dims = rng.integers(1, 3, 200)
    vals = rng.standard_normal(200)
    kn, kd, kv = meshgen.last_write_wins(nodes, dims, vals)
    want = {}
    for n_, d_, v_ in zip(nodes.tolist(), dims.tolist(), vals.tolist()):
        want[(n_, d_)] = +v_
    assert list(zip(kn.tolist(), kd.tolist())) == list(want.keys())
    assert kv.tolist() == list(want.values())
