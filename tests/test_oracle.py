"""The CPU oracle pinned against closed forms, the reference's NumPy transcriptions and golden values,
and brute force on small inputs. Each test names the reference test or formula it restates."""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import kuhn_grid


def dense_csr(ptr, col, val, n_cols):
    rows = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
    out = np.zeros((len(ptr) - 1, n_cols))
    out[rows, col] = val
    return out


# ------------------------------------------------------------------------------------------------
# Gradients and assembly
# ------------------------------------------------------------------------------------------------

def test_gradients_of_the_reference_tet(oracle_mod):
    """tests/test_solver.py:18-31: unit right tet in mm -> g in 1/m, vol = 1e-9/6 m^3."""
    nodes = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    g, vol = oracle_mod.p1_gradients(nodes, np.array([[0, 1, 2, 3]], np.int32))
    np.testing.assert_allclose(g[0], [[-1e3, -1e3, -1e3], [1e3, 0, 0], [0, 1e3, 0], [0, 0, 1e3]], rtol=1e-14)
    np.testing.assert_allclose(vol, [1e-9 / 6], rtol=1e-14)


def test_gradients_reproduce_linear_functions(oracle_mod):
    """P1 gradients are those of the barycentric coordinates: sum_i g_i = 0 and
    sum_i x_i (x) g_i = I for the node coordinates x in metres; vol = |det| / 6."""
    nodes, tets = kuhn_grid(3, 3, 3, spacing=7.0, jitter=0.2, seed=4)
    g, vol = oracle_mod.p1_gradients(nodes, tets)
    np.testing.assert_allclose(g.sum(axis=1), 0.0, atol=1e-9)
    x = nodes[tets] * 1e-3
    ident = np.einsum("mij,mik->mjk", x, g)
    np.testing.assert_allclose(ident, np.broadcast_to(np.eye(3), ident.shape), atol=1e-11)
    p = nodes[tets]
    det = np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0])
    np.testing.assert_allclose(vol, np.abs(det) / 6 * 1e-9, rtol=1e-12)


@pytest.fixture(scope="module")
def small_system(oracle_mod):
    nodes, tets = kuhn_grid(4, 3, 3, spacing=9.0, jitter=0.15, seed=7)
    rng = np.random.default_rng(1)
    tags = rng.integers(1, 6, tets.shape[0]).astype(np.int32)
    n = nodes.shape[0]
    g, vol = oracle_mod.p1_gradients(nodes, tets)
    cond = oracle_mod.conductivity_per_tet(tags)
    ptr, idx = oracle_mod.node2corner(tets, n)
    indptr, indices = oracle_mod.stiffness_pattern(tets, ptr, idx, n)
    data = oracle_mod.stiffness_values(g, vol * cond, tets, ptr, idx, indptr, indices)
    return dict(nodes=nodes, tets=tets, tags=tags, g=g, vol=vol, cond=cond, ptr=ptr, idx=idx,
                indptr=indptr, indices=indices, data=data, n=n)


def test_conductivity_table_and_unknown_tags(oracle_mod):
    """fem/assembly.py:15-42: the table values; an unknown tag raises."""
    np.testing.assert_allclose(oracle_mod.conductivity_per_tet(np.array([1, 2, 3, 4, 5], np.int32)),
                               [0.126, 0.275, 1.654, 0.010, 0.465])
    with pytest.raises(ValueError):
        oracle_mod.conductivity_per_tet(np.array([1, 42], np.int32))


def test_stiffness_equals_brute_force_dense_assembly(small_system):
    """A = sum_e vol_e sigma_e g_e g_e^T, scattered densely (tests/test_solver.py:55-69 checks the
    same matrix through its symmetry and row sums)."""
    s = small_system
    n = s["n"]
    dense = np.zeros((n, n))
    local = np.einsum("m,mik,mjk->mij", s["vol"] * s["cond"], s["g"], s["g"])
    t = s["tets"]
    np.add.at(dense, (t[:, :, None], t[:, None, :]), local)
    got = dense_csr(s["indptr"], s["indices"], s["data"], n)
    np.testing.assert_allclose(got, dense, rtol=1e-12, atol=1e-12 * np.abs(dense).max())
    np.testing.assert_allclose(got, got.T, rtol=0, atol=1e-15 * np.abs(dense).max())
    assert np.abs(got.sum(axis=1)).max() <= 1e-12 * np.abs(dense).max()


def test_pattern_is_the_set_of_tet_sharing_pairs(small_system):
    """solver/fem/pattern.cu:65-87,144-299: sorted, distinct, exactly the node pairs of some tet."""
    s = small_system
    t = s["tets"]
    pairs = set(zip(np.repeat(t, 4, axis=1).ravel().tolist(), np.tile(t, (1, 4)).ravel().tolist()))
    rows = np.repeat(np.arange(s["n"]), np.diff(s["indptr"]))
    got = set(zip(rows.tolist(), s["indices"].tolist()))
    assert got == pairs
    for i in range(s["n"]):
        seg = s["indices"][s["indptr"][i]:s["indptr"][i + 1]]
        assert np.all(np.diff(seg) > 0)


def test_node2corner_is_the_transpose_of_the_connectivity(small_system):
    """tests/test_solver.py:72-82: corner c = 4 e + i is listed under node tets[e, i], ascending."""
    s = small_system
    flat = s["tets"].ravel()
    for node in range(s["n"]):
        seg = s["idx"][s["ptr"][node]:s["ptr"][node + 1]]
        assert np.all(flat[seg] == node)
        assert np.all(np.diff(seg) > 0)
    assert s["ptr"][-1] == flat.size


def test_grounded_operator_is_spd(small_system, oracle_mod):
    """tests/test_realmesh.py:95-112: dropping the ground row/column leaves an SPD matrix."""
    s = small_system
    ground = oracle_mod.ground_node_of(s["nodes"])
    rp, rc, rv = oracle_mod.reduce_csr(s["indptr"], s["indices"], s["data"], ground)
    a = dense_csr(rp, rc, rv, s["n"] - 1)
    full = dense_csr(s["indptr"], s["indices"], s["data"], s["n"])
    keep = np.arange(s["n"]) != ground
    np.testing.assert_array_equal(a, full[np.ix_(keep, keep)])
    assert np.linalg.eigvalsh(a).min() > 0


def test_ground_node_is_lowest_z_with_lowest_index_on_ties(oracle_mod):
    """fem/solve.py:287-289, tests/test_solver.py:95-138."""
    nodes = np.array([[0, 0, 5.0], [1, 0, -2.0], [2, 0, -2.0], [3, 0, 1.0]])
    assert oracle_mod.ground_node_of(nodes) == 1


# ------------------------------------------------------------------------------------------------
# Mesh order
# ------------------------------------------------------------------------------------------------

def _spread_bits(v: int) -> int:
    out = 0
    for i in range(21):
        out |= ((v >> i) & 1) << (3 * i)
    return out


def test_morton_keys_follow_the_bit_interleave_definition(oracle_mod):
    """fem/solve.py:299-323: 21-bit quantisation over the bounding box, x in bit 0, y 1, z 2."""
    rng = np.random.default_rng(3)
    pts = rng.uniform(-90, 90, (200, 3))
    pts[5] = pts[6]  # a duplicate
    lo = pts.min(0)
    span = np.maximum(pts.max(0) - lo, 1e-12)
    q = ((pts - lo) / span * float((1 << 21) - 1)).astype(np.uint64)
    want = [sum(_spread_bits(int(q[i, a])) << a for a in range(3)) for i in range(len(pts))]
    np.testing.assert_array_equal(oracle_mod.morton_keys(pts), np.array(want, dtype=np.uint64))


def test_spatial_order_is_idempotent(head8, oracle_mod):
    """fem/solve.py:333-348: a mesh already in Morton order sorts to the identity."""
    nodes, tets, tags, skin, node_perm, tet_perm = oracle_mod.reorder_mesh(
        head8.nodes_mm, head8.tet_nodes, head8.tet_tags, head8.skin_tris)
    assert not np.array_equal(node_perm, np.arange(len(node_perm)))
    again = oracle_mod.spatial_order(nodes, tets)
    np.testing.assert_array_equal(again[0], np.arange(nodes.shape[0]))
    np.testing.assert_array_equal(again[1], np.arange(tets.shape[0]))
    # Tets ordered by their lowest renumbered node.
    assert np.all(np.diff(tets.min(axis=1)) >= 0)


# ------------------------------------------------------------------------------------------------
# AMG
# ------------------------------------------------------------------------------------------------

@pytest.fixture(scope="module")
def head8_ctx(head8, oracle_mod):
    return oracle_mod.build_context(head8.nodes_mm, head8.tet_nodes, head8.tet_tags, head8.skin_tris)


def test_l1_dinv_formula_and_spectral_bound(head8_ctx, oracle_mod):
    """solver/amg/l1.cu:15-35; tests/test_solver.py:141-162: rho(D^-1 A) <= 1."""
    a = head8_ctx.hier.levels[0].a
    dinv = oracle_mod.l1_dinv(a)
    d = a.dense()
    want = 1.0 / (np.sign(np.diag(d)) * np.abs(d).sum(axis=1))
    np.testing.assert_allclose(dinv, want.astype(np.float32), rtol=2e-7)
    # D^-1 A is similar to the symmetric D^-1/2 A D^-1/2 (dinv > 0 here).
    assert np.all(dinv > 0)
    h = np.sqrt(dinv.astype(np.float64))
    s = sp.csr_matrix(h[:, None] * d * h[None, :])
    lam_max = spla.eigsh(s, k=1, which="LA", return_eigenvectors=False, tol=1e-8)[0]
    assert lam_max <= 1.0 + 1e-6


def test_size4_aggregates_partition_into_connected_sets(head8_ctx, oracle_mod):
    """tests/test_solver_kernels.py:454-473: every row in exactly one aggregate, labels dense;
    each aggregate is connected in the matrix graph; repeated calls agree (:578-612)."""
    a = head8_ctx.hier.levels[0].a
    agg, nc = oracle_mod.select_size4(a)
    assert agg.min() == 0 and agg.max() == nc - 1
    assert np.all(np.bincount(agg, minlength=nc) >= 1)
    agg2, nc2 = oracle_mod.select_size4(a)
    np.testing.assert_array_equal(agg, agg2)
    g = sp.csr_matrix((np.ones_like(a.val), a.col, a.ptr), shape=(a.n, a.n))
    for c in range(0, nc, max(1, nc // 200)):
        members = np.nonzero(agg == c)[0]
        sub = g[members][:, members]
        ncomp, _ = sp.csgraph.connected_components(sub, directed=False)
        assert ncomp == 1


def test_spgemm_and_transpose_match_dense_products(oracle_mod):
    rng = np.random.default_rng(5)
    a = sp.random(60, 40, density=0.1, random_state=1, format="csr", dtype=np.float32)
    b = sp.random(40, 50, density=0.1, random_state=2, format="csr", dtype=np.float32)
    a.sort_indices(); b.sort_indices()
    ca = oracle_mod.Csr(a.indptr.astype(np.int32), a.indices.astype(np.int32), a.data, 40)
    cb = oracle_mod.Csr(b.indptr.astype(np.int32), b.indices.astype(np.int32), b.data, 50)
    c = oracle_mod.spgemm(ca, cb)
    np.testing.assert_allclose(c.dense(), a.toarray().astype(np.float64) @ b.toarray(), rtol=1e-5, atol=1e-6)
    for i in range(c.n):
        assert np.all(np.diff(c.col[c.ptr[i]:c.ptr[i + 1]]) > 0)
    t = oracle_mod.transpose(ca)
    np.testing.assert_array_equal(t.dense(), a.toarray().T)
    del rng


def test_prolongators_and_galerkin_match_their_formulas(head8_ctx, oracle_mod):
    """fem/solve.py:94-134: P_tent^T P_tent = I; P = (I - 4/3 D^-1 A) P_tent; A_c = P^T A P."""
    lv = head8_ctx.hier.levels[0]
    nc = int(lv.agg.max()) + 1
    pt = oracle_mod.tentative_prolongator(lv.agg, nc)
    ptd = pt.dense()
    np.testing.assert_allclose(ptd.T @ ptd, np.eye(nc), atol=1e-6)
    a = lv.a.dense()
    want_p = (np.eye(a.shape[0]) - (4.0 / 3.0) * np.diag(lv.dinv.astype(np.float64)) @ a) @ ptd
    np.testing.assert_allclose(lv.p.dense(), want_p, atol=1e-6)
    ac = oracle_mod.galerkin(lv.a, lv.p)
    want_ac = lv.p.dense().T @ a @ lv.p.dense()
    np.testing.assert_allclose(ac.dense(), want_ac, rtol=1e-4, atol=1e-5 * np.abs(want_ac).max())
    np.testing.assert_allclose(lv.r.dense(), lv.p.dense().T)


def test_chebyshev_coefficients_give_the_chebyshev_error_polynomial(oracle_mod):
    """solver/amg/vcycle.cu:434-463: after d steps from x = 0 the error factor at eigenvalue lam of
    D^-1 A is T_d((theta - lam)/delta) / T_d(theta/delta) over [1/12, 1]."""
    for degree in (1, 2, 3, 4, 6):
        ch = oracle_mod.chebyshev_coefficients(degree, 12.0)
        lo, hi = 1.0 / 12.0, 1.0
        theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
        lam = np.linspace(0.0, 1.0, 41)
        x_prev = np.zeros_like(lam)
        x = float(ch["alpha0"]) * np.ones_like(lam)  # b = 1
        for k in range(degree - 1):
            x_next = float(ch["c_cur"][k]) * x + float(ch["c_prev"][k]) * x_prev + float(ch["alpha"][k]) * (1 - lam * x)
            x_prev, x = x, x_next
        err = 1 - lam * x
        tch = np.polynomial.chebyshev.Chebyshev.basis(degree)
        want = tch((theta - lam) / delta) / tch(theta / delta)
        np.testing.assert_allclose(err, want, atol=2e-6)


def _chebyshev_error_matrix(a, dinv, degree, ratio):
    """T_d(D^-1 A) for the scaled Chebyshev polynomial T_d((theta - x)/delta) / T_d(theta/delta) on
    [1/ratio, 1], from the eigenvectors of the symmetric D^-1/2 A D^-1/2 (independent of the
    smoother's three-term recurrence)."""
    lo, hi = 1.0 / ratio, 1.0
    theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
    d = np.sqrt(dinv)
    lam, vec = np.linalg.eigh(d[:, None] * a * d[None, :])
    t = np.polynomial.chebyshev.Chebyshev.basis(degree)
    f = t((theta - lam) / delta) / t(theta / delta)
    return (vec * f) @ vec.T * d[:, None] / d[None, :]


def _vcycle_closed_form(h, level, degree, ratio, coarse=True):
    """Dense V-cycle operator from its definition: error propagation E (I - P M_c P^T A) E with E
    the smoother's error polynomial, so M = (I - E (I - P M_c P^T A) E) A^-1, recursing down to the
    exact coarsest inverse."""
    if level == len(h.levels):
        return np.linalg.inv(h.coarse.dense().astype(np.float64))
    lv = h.levels[level]
    a = lv.a.dense().astype(np.float64)
    e = _chebyshev_error_matrix(a, lv.dinv.astype(np.float64), degree, ratio)
    p = lv.p.dense().astype(np.float64)
    corr = p @ _vcycle_closed_form(h, level + 1, degree, ratio) @ p.T @ a if coarse else 0.0
    eye = np.eye(a.shape[0])
    return (eye - e @ (eye - corr) @ e) @ np.linalg.inv(a)


def test_vcycle_equals_its_dense_closed_form(oracle_mod, monkeypatch):
    """solver/amg/vcycle.cu:560-648 composed from dense matrices on a two-level hierarchy (a
    1,728-node jittered two-tissue box, coarse rows floor lowered to 16 so that two smoothed levels
    form): the oracle's V-cycle applied to every unit vector equals the closed form to the FP32
    rounding of the Chebyshev coefficients. One smoothing step more or the coarse correction left
    out moves the operator by far more than the tolerance."""
    nodes, tets = kuhn_grid(11, 11, 11, spacing=5.0, jitter=0.1, seed=3)
    tags = np.ones(len(tets), np.int32)
    tags[: len(tets) // 3] = 2
    from conftest import boundary_tris
    ctx = oracle_mod.build_context(nodes, tets, tags, boundary_tris(tets, nodes), hierarchy=False)
    monkeypatch.setattr(oracle_mod, "MIN_COARSE_ROWS", 16)
    h = oracle_mod.build_hierarchy(ctx.reduced)
    assert len(h.levels) == 2 and h.coarse.n >= 16
    n = h.levels[0].a.n
    got = np.stack([oracle_mod.vcycle_apply(h, e) for e in np.eye(n)], axis=1)
    want = _vcycle_closed_form(h, 0, 2, 12.0)
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 2e-6 * scale
    assert np.abs(got - _vcycle_closed_form(h, 0, 3, 12.0)).max() > 1e-2 * scale
    assert np.abs(got - _vcycle_closed_form(h, 0, 2, 12.0, coarse=False)).max() > 1e-2 * scale


def test_vcycle_is_symmetric_and_positive(head8_ctx, oracle_mod):
    """A symmetric V-cycle (same polynomial smoother before and after, R = P^T) is an SPD
    preconditioner: u^T M v = v^T M u and v^T M v > 0."""
    h = head8_ctx.hier
    n = h.levels[0].a.n
    rng = np.random.default_rng(11)
    u, v = rng.standard_normal(n), rng.standard_normal(n)
    mu, mv = oracle_mod.vcycle_apply(h, u), oracle_mod.vcycle_apply(h, v)
    assert abs(u @ mv - v @ mu) <= 1e-10 * abs(u @ mv)
    assert v @ mv > 0


def test_pcg_matches_a_direct_solve(head8_ctx, oracle_mod):
    """tests/test_solver.py:165-188: the AMG-PCG answer equals a direct FP64 solve."""
    c = head8_ctx
    a = sp.csr_matrix((c.reduced_val64, c.reduced.col, c.reduced.ptr), shape=(c.reduced.n,) * 2)
    rng = np.random.default_rng(2)
    b = rng.standard_normal(c.reduced.n)
    x, iters, rel = oracle_mod.pcg(c.hier, c.reduced, c.reduced_val64, b, 1e-12)
    want = spla.spsolve(a.tocsc(), b)
    assert rel <= 1e-12 and iters < 200
    np.testing.assert_allclose(x, want, rtol=0, atol=1e-9 * np.abs(want).max())


# ------------------------------------------------------------------------------------------------
# Coil, RHS, reconstruction
# ------------------------------------------------------------------------------------------------

def dadt_reference(s, m, r, didt):
    """tests/test_placement.py:51-59: dA/dt(r) = didt mu0/4pi sum_j m_j x (r - s_j)/|r - s_j|^3."""
    diff = r[:, None, :] - s[None, :, :]
    dist3 = np.linalg.norm(diff, axis=2)[:, :, None] ** 3
    return didt * 1e-7 * (np.cross(m[None, :, :], diff) / dist3).sum(axis=1)


def test_dadt_matches_the_closed_form_dipole_sum(oracle_mod):
    rng = np.random.default_rng(8)
    s = rng.normal(0, 0.02, (37, 3))
    m = rng.normal(0, 1e-3, (37, 3))
    r = rng.normal(0, 0.05, (101, 3)) + [0, 0, -0.08]
    np.testing.assert_allclose(oracle_mod.dadt_direct(s, m, r, 1e6), dadt_reference(s, m, r, 1e6), rtol=1e-12)


def test_dadt_at_nodes_applies_the_frame(oracle_mod, coil):
    """fem/placement.py:117-155: dipoles are moved into the frame, moments rotated, not translated."""
    rng = np.random.default_rng(9)
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    q *= np.sign(np.linalg.det(q))
    tf = np.eye(4)
    tf[:3, :3] = q
    tf[:3, 3] = [5.0, -12.0, 100.0]
    nodes = rng.uniform(-40, 40, (50, 3))
    got = oracle_mod.dadt_at_nodes(coil.positions_m, coil.moments, tf, 1e6, nodes)
    s = (coil.positions_m * 1e3 @ q.T + tf[:3, 3]) * 1e-3
    m = coil.moments @ q.T
    np.testing.assert_allclose(got, dadt_reference(s, m, nodes * 1e-3, 1e6), rtol=1e-11)


def closest_on_tri(p, a, b, c):
    """Ericson's closest point on a triangle (Real-Time Collision Detection, 5.1.5), the reference's
    test transcription (tests/test_placement.py:21-46) restated."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + d1 / (d1 - d3) * ab
    cp_ = p - c
    d5, d6 = ab @ cp_, ac @ cp_
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + d2 / (d2 - d6) * ac
    va = d3 * d6 - d5 * d4
    if va <= 0 and d4 - d3 >= 0 and d5 - d6 >= 0:
        return b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b)
    den = 1.0 / (va + vb + vc)
    return a + vb * den * ab + vc * den * ac


def test_placement_frame_on_a_flat_scalp(oracle_mod):
    """solver/coil/place.cu:153-230: c = projection + d n, z = -n, y = in-plane handle, x = y x z."""
    nodes = np.array([[-100.0, -100, 0], [100, -100, 0], [100, 100, 0], [-100, 100, 0]])
    tris = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    normals = oracle_mod.skin_normals(nodes, tris)
    np.testing.assert_allclose(normals, [[0, 0, 1], [0, 0, 1]], atol=1e-15)
    tf, deg = oracle_mod.place_transforms([10.0, 20.0, 7.0], [10.0, 70.0, 9.0], [4.0], nodes, tris, normals)
    assert deg[0] == 0
    np.testing.assert_allclose(tf[0, :3, 3], [10, 20, 4], atol=1e-12)
    np.testing.assert_allclose(tf[0, :3, 2], [0, 0, -1], atol=1e-12)
    np.testing.assert_allclose(tf[0, :3, 1], [0, 1, 0], atol=1e-12)
    np.testing.assert_allclose(tf[0, :3, 0], [-1, 0, 0], atol=1e-12)
    np.testing.assert_allclose(tf[0, 3], [0, 0, 0, 1], atol=0)
    _, deg = oracle_mod.place_transforms([10.0, 20.0, 7.0], [10.0, 20.0, 50.0], [4.0], nodes, tris, normals)
    assert deg[0] == 1  # handle along the normal


def test_placement_projects_to_the_brute_force_closest_point(head8, oracle_mod):
    """tests/test_placement.py:264-273: projection = nearest skin point over all triangles."""
    nodes, tets, tags, skin, _, _ = oracle_mod.reorder_mesh(head8.nodes_mm, head8.tet_nodes, head8.tet_tags,
                                                            head8.skin_tris)
    normals = oracle_mod.skin_normals(nodes, skin)
    rng = np.random.default_rng(4)
    for _ in range(3):
        d = rng.standard_normal(3)
        d[2] = abs(d[2]) + 0.5
        p = d / np.linalg.norm(d) * 110.0
        tf, deg = oracle_mod.place_transforms(p, None, [0.0], nodes, skin, normals)
        tri = nodes[skin]
        best = min((np.sum((closest_on_tri(p, *t) - p) ** 2), i) for i, t in enumerate(tri))
        q = closest_on_tri(p, *tri[best[1]])
        np.testing.assert_allclose(tf[0, :3, 3], q, atol=1e-9)
        rot = tf[0, :3, :3]
        np.testing.assert_allclose(rot.T @ rot, np.eye(3), atol=1e-12)
        assert np.linalg.det(rot) == pytest.approx(1.0)


def test_skin_normals_point_outward_on_the_sphere_head(head8, oracle_mod):
    normals = oracle_mod.skin_normals(head8.nodes_mm, head8.skin_tris)
    centres = head8.nodes_mm[head8.skin_tris].mean(axis=1)
    radial = centres / np.linalg.norm(centres, axis=1, keepdims=True)
    np.testing.assert_allclose(np.linalg.norm(normals, axis=1), 1.0, rtol=1e-12)
    assert np.einsum("ij,ij->i", normals, radial).min() > 0.8


def test_rhs_matches_the_corner_transcription(small_system, oracle_mod):
    """tests/test_solver_kernels.py:89-99: b[node] = sum over its corners of neg_vc (dadt_e . g_c)."""
    s = small_system
    rng = np.random.default_rng(6)
    dadt = rng.standard_normal((s["tets"].shape[0], 3))
    neg_vc = -(s["vol"] * s["cond"])
    want = np.zeros(s["n"])
    np.add.at(want, s["tets"].ravel(), (neg_vc[:, None] * np.einsum("mk,mik->mi", dadt, s["g"])).ravel())
    np.testing.assert_allclose(oracle_mod.rhs(dadt, s["g"], neg_vc, s["ptr"], s["idx"]), want, rtol=1e-12,
                               atol=1e-14 * np.abs(want).max())


def test_reconstruct_is_exact_for_a_linear_potential(small_system, oracle_mod):
    """solver/fem/reconstruct.cu:12-31: E = -grad v - dA/dt; v = a . x gives E = -a - dA/dt."""
    s = small_system
    a = np.array([3.0, -1.0, 2.0])
    v = (s["nodes"] * 1e-3) @ a
    dadt = np.tile([0.5, 0.25, -1.0], (s["tets"].shape[0], 1))
    e, mag = oracle_mod.reconstruct(v, s["tets"], s["g"], dadt)
    np.testing.assert_allclose(e, np.broadcast_to(-a - dadt[0], e.shape), atol=1e-10)
    np.testing.assert_allclose(mag, np.linalg.norm(e, axis=1), rtol=1e-14)


def test_uniform_induced_field_in_a_homogeneous_conductor_gives_zero_E(oracle_mod):
    """Physics closed form: with uniform dA/dt = A' in a homogeneous body the Neumann problem is
    solved by v = -A'.x, so E = -grad v - A' = 0 everywhere. Exercises RHS, grounding, AMG-PCG and
    reconstruction together."""
    nodes, tets = kuhn_grid(6, 6, 5, spacing=6.0, jitter=0.1, seed=3)
    ctx = oracle_mod.build_context(nodes, tets, np.full(tets.shape[0], 2, np.int32), np.empty((0, 3), np.int32))
    ap = np.array([1.0, -2.0, 0.5])
    dadt = np.tile(ap, (ctx.tet_nodes.shape[0], 1))
    b = oracle_mod.rhs(dadt, ctx.g, -(ctx.vols * ctx.cond), ctx.n2c_ptr, ctx.n2c_idx)
    keep = np.arange(nodes.shape[0]) != ctx.ground
    x, _, rel = oracle_mod.pcg(ctx.hier, ctx.reduced, ctx.reduced_val64, b[keep], 1e-13)
    v = np.zeros(nodes.shape[0])
    v[keep] = x
    np.testing.assert_allclose(v - v[ctx.ground], -(ctx.nodes_mm - ctx.nodes_mm[ctx.ground]) * 1e-3 @ ap,
                               atol=1e-9)
    _, mag = oracle_mod.reconstruct(v, ctx.tet_nodes, ctx.g, dadt)
    assert mag.max() <= 1e-8 * np.linalg.norm(ap)


# ------------------------------------------------------------------------------------------------
# Metrics (golden values of the reference's tests/test_metrics.py)
# ------------------------------------------------------------------------------------------------

MAGNE = np.array([1.0, 2.0, 3.0, 100.0])
VOLS = np.ones(4)
BARY = np.array([[0.0, 0, 0], [1, 0, 0], [2, 0, 0], [9, 9, 9]])
GM = np.array([True, True, True, False])


def test_metrics_golden_values(oracle_mod):
    """tests/test_metrics.py:9-80: GM of [1, 2, 3 | 100] with unit volumes."""
    s = oracle_mod.summarize(MAGNE[GM], VOLS[GM], BARY[GM])
    assert s["peak_magnE"] == 3.0
    np.testing.assert_allclose(s["peak_location_mm"], [2, 0, 0])
    assert s["region_volume_m3"] == 3.0
    assert s["focality_m3"]["0.5"] == 2.0  # p99.9 saturates to 3: elements >= 1.5
    assert s["distribution"]["mean"] == pytest.approx(2.0)
    assert s["distribution"]["std"] == pytest.approx(np.sqrt(2 / 3))
    assert s["distribution"]["p50"] == pytest.approx(2.0)
    np.testing.assert_allclose(s["center_of_gravity_mm"], [(0 + 2 + 6) / 6, 0, 0])


def test_weighted_quantiles_reduce_to_hazen(oracle_mod):
    """tests/test_metrics.py:89-96."""
    rng = np.random.default_rng(0)
    values = rng.random(1001)
    qs = np.array([0.01, 0.1, 0.5, 0.95, 0.99, 0.999])
    np.testing.assert_allclose(oracle_mod.weighted_quantiles(values, np.ones_like(values), qs),
                               np.quantile(values, qs, method="hazen"), atol=1e-9)


def test_weighted_quantiles_weights_and_scale(oracle_mod):
    """tests/test_metrics.py:99-119."""
    v = np.array([1.0, 2.0])
    assert oracle_mod.weighted_quantiles(v, np.array([1.0, 9.0]), np.array([0.5]))[0] > 1.8
    assert oracle_mod.weighted_quantiles(v, np.array([9.0, 1.0]), np.array([0.5]))[0] < 1.2
    rng = np.random.default_rng(1)
    values, weights = rng.random(50), rng.random(50) + 0.1
    qs = np.array([0.25, 0.5, 0.9])
    np.testing.assert_allclose(oracle_mod.weighted_quantiles(values, weights, qs),
                               oracle_mod.weighted_quantiles(values, weights * 1e6, qs), rtol=1e-12)


def test_focality_anchors_to_the_percentile_not_the_max(oracle_mod):
    """tests/test_metrics.py:38-52: one hot sliver must not drag the threshold up."""
    magn = np.concatenate([np.ones(9999), [100.0]])
    s = oracle_mod.summarize(magn, np.ones(10000), np.zeros((10000, 3)))
    assert s["peak_magnE"] == 100.0
    assert s["distribution"]["p99.9"] == pytest.approx(1.0)
    assert s["focality_m3"]["0.5"] == pytest.approx(10000.0)


def test_single_element_distribution(oracle_mod):
    """tests/test_metrics.py:160-168."""
    s = oracle_mod.summarize(np.array([7.0]), np.array([3.0]), np.zeros((1, 3)))
    assert s["distribution"]["mean"] == pytest.approx(7.0)
    assert s["distribution"]["std"] == pytest.approx(0.0)
    assert s["distribution"]["p50"] == pytest.approx(7.0)
