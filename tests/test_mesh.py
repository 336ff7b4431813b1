"""Host mesh work: reader/writer, generators, geometry, RCM order, face colouring (CPU only).

Integer work is checked bit-exactly; the FP64 geometry is checked bit-exactly against the oracle
(same formulas, same order) and against closed forms."""
import os

import numpy as np
import pytest

import mstcfd as M
from helpers import brick_grid, periodic_in_x, quad_grid
from oracle import pyoracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_MSH = "/root/reference/MST-CFD/msh"
FIXTURES = ["2d-shockwavepipe-2", "2d-stairW-1", "2d-stair-st-2", "2d-stair-un-5-tri"]

# Appendix-B facts of all 8 bundled meshes: cells, nodes, faces, interior faces, BC types present
BUNDLED = {
    "2d-shockwavepipe-2": (18282, 9382, 27663, 27183, {3}),
    "2d-stair-st-2": (16700, 17172, 33871, 32929, {10, 5, 3}),
    "2d-stair-un-3-loose-tri": (3111, 1691, 4801, 4532, {10, 5, 3}),
    "2d-stair-un-3-tri": (6094, 3236, 9329, 8953, {10, 5, 3}),
    "2d-stair-un-4-tri": (4853, 2595, 7447, 7112, {10, 5, 3}),
    "2d-stair-un-5-tri": (3103, 1687, 4789, 4520, {10, 5, 3}),
    "2d-stairW-1": (19594, 10048, 29641, 29141, {10, 7, 3, 5}),
    "2d-stairW-2-st": (34000, 34501, 68500, 67500, {10, 3, 5}),
}


def fixture(name):
    d = dict(np.load(os.path.join(GOLD, f"msh_{name}.npz")))
    for k in ("dim", "n_cells", "n_int"):
        d[k] = int(d[k])
    return d


@pytest.mark.skipif(not os.path.isdir(REF_MSH), reason="reference meshes not present")
@pytest.mark.parametrize("name", sorted(BUNDLED))
def test_reader_on_bundled_meshes(name):
    m = M.Mesh.read_msh(os.path.join(REF_MSH, name + ".msh"))
    nc, nn, nf, ni, types = BUNDLED[name]
    assert (m.n_cells, m.n_nodes, m.n_faces, m.n_int) == (nc, nn, nf, ni)
    raw = m.raw()
    assert set(np.unique(raw["ftype"][raw["c1"] < 0]).tolist()) == types
    geo = M.Prep(m).geometry()
    assert (geo["dac"] == 1).all()  # every face is "branch 2" (area vector out of c0)
    if name in FIXTURES:
        fx = fixture(name)
        for k in ("xyz", "fnodes", "c0", "c1", "ftype"):
            assert np.array_equal(raw[k], fx[k]), k
        assert np.array_equal(m.zones()[:, 1:], fx["zones"][:, 1:])


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_round_trip_and_geometry(name, tmp_path):
    fx = fixture(name)
    m = M.Mesh.from_raw(fx)
    assert (m.n_cells, m.n_int) == (fx["n_cells"], fx["n_int"])
    p = tmp_path / "rt.msh"
    m.write_msh(p)
    m2 = M.Mesh.read_msh(p)
    r2 = m2.raw()
    for k in ("xyz", "fnodes", "c0", "c1", "ftype"):
        assert np.array_equal(r2[k], fx[k]), k
    g = M.Prep(m).geometry()
    go = O.OracleMesh(fx).geometry()
    for k in g:
        assert np.array_equal(g[k], go[k], equal_nan=True), k
    area = {"2d-shockwavepipe-2": 0.2, "2d-stairW-1": None}.get(name)
    if area:
        assert abs(g["vol"].sum() - area) < 1e-6  # single-precision node coordinates in the file


def test_generated_counts_match_configs():
    # C1: 2000x2 quads -> 8000 triangles on [0,1]x[0,0.01], walls
    m = M.Mesh.gen_tri(2000, 2, 0, 1, 0, 0.01, 0.1, 12345, (M.BC_WALL,) * 4)
    assert (m.n_cells, m.n_faces, m.n_int, m.n_nodes) == (8000, 14002, 9998, 6003)
    # C2: 1000x1000 quads -> 2M triangles, periodic: 2,998,000 interior + 4000 periodic faces
    m = M.Mesh.gen_tri(1000, 1000, -10, 10, -10, 10, 0.1, 12345, (M.BC_PERIODIC,) * 4)
    assert (m.n_cells, m.n_faces, m.n_int, m.n_nodes) == (2000000, 3002000, 2998000, 1002001)
    p = M.Prep(m)
    assert p.n_periodic == 2000 and p.maxf == 3  # 4000 periodic faces form 2000 pairs
    # C3 construction at n=8: 6 n^3 tets, 12 n^3 - 6 n^2 interior faces, 12 n^2 boundary faces
    n = 8
    m = M.Mesh.gen_tet(n, n, n, (0, 0, 0), (1, 1, 1), 0.15, 12345, (M.BC_INLET, M.BC_OUTLET) + (M.BC_SYMMETRY,) * 4)
    assert (m.n_cells, m.n_int, m.n_faces, m.n_nodes) == (6 * n ** 3, 12 * n ** 3 - 6 * n ** 2,
                                                         12 * n ** 3 + 6 * n ** 2, (n + 1) ** 3)


@pytest.mark.parametrize("dim", [2, 3])
def test_generated_geometry_closed_forms(dim):
    if dim == 2:
        m = M.Mesh.gen_tri(30, 20, 0, 3, -1, 1, 0.2, 5, (M.BC_WALL, M.BC_OUTLET, M.BC_SYMMETRY, M.BC_INLET))
        area = 3 * 2
    else:
        m = M.Mesh.gen_tet(5, 4, 3, (0, 0, 0), (1, 2, 3), 0.2, 5, (M.BC_WALL,) * 6)
        area = 6
    raw = m.raw()
    g = M.Prep(m).geometry()
    go = O.OracleMesh(raw).geometry()
    for k in g:
        assert np.array_equal(g[k], go[k]), k
    assert (g["dac"] == -1).all()  # generated faces point into c0
    assert abs(g["vol"].sum() - area) <= 1e-12 * area
    # closed cells: sum of outward area vectors vanishes
    n = g["S"] * g["dac"][:, None]
    acc = np.zeros((m.n_cells, dim))
    np.add.at(acc, raw["c0"], n)
    mk = raw["c1"] >= 0
    np.add.at(acc, raw["c1"][mk], -n[mk])
    mag = np.zeros(m.n_cells)
    np.add.at(mag, raw["c0"], np.abs(n).sum(1))
    np.add.at(mag, raw["c1"][mk], np.abs(n[mk]).sum(1))
    assert (np.abs(acc).max(1) <= 1e-14 * mag).all()
    if dim == 2:  # Heron (the reference's area) against the shoelace formula
        tri = [[] for _ in range(m.n_cells)]
        for f in range(m.n_faces):
            for c in (raw["c0"][f], raw["c1"][f]):
                if c >= 0:
                    tri[c] += list(raw["fnodes"][f])
        x = raw["xyz"]
        shoe = np.array([0.5 * abs(np.cross(x[sorted(set(t))[1]] - x[sorted(set(t))[0]],
                                            x[sorted(set(t))[2]] - x[sorted(set(t))[0]])) for t in tri])
        assert np.abs(g["vol"] - shoe).max() <= 1e-13 * shoe.max()
        # cell centre = centroid (mean of face midpoints == mean of vertices for triangles)
        cen = np.array([x[sorted(set(t))].mean(0) for t in tri])
        assert np.abs(g["ccen"] - cen).max() <= 1e-13


def test_jitter_bounds_and_periodic_match():
    nx, ny, j = 20, 10, 0.1
    m = M.Mesh.gen_tri(nx, ny, -10, 10, -5, 5, j, 9, (M.BC_PERIODIC,) * 4)
    x = m.raw()["xyz"].reshape(ny + 1, nx + 1, 2)
    hx, hy = 20 / nx, 10 / ny
    gx = -10 + hx * np.arange(nx + 1)
    gy = -5 + hy * np.arange(ny + 1)
    assert np.abs(x[:, :, 0] - gx[None, :]).max() <= j * hx + 1e-12
    assert np.abs(x[:, :, 1] - gy[:, None]).max() <= j * hy + 1e-12
    assert np.array_equal(x[:, 0, 1], x[:, -1, 1]) and np.array_equal(x[0, :, 0], x[-1, :, 0])
    assert (x[:, 0, 0] == -10).all() and (x[0, :, 1] == -5).all()
    raw = m.raw()
    g = M.Prep(m).geometry()
    p = raw["partner"]
    per = np.where(p >= 0)[0]
    assert np.array_equal(p[p[per]], per)
    assert np.array_equal(g["S"][per], -g["S"][p[per]])  # exactly opposite area vectors


@pytest.fixture
def cell_order():
    """Sets the process-wide cell order for one test and restores it afterwards."""
    old = M.get_cell_order()
    yield M.set_cell_order
    M.set_cell_order(old)


@pytest.mark.parametrize("order", [M.ORDER_RCM, M.ORDER_HILBERT])
@pytest.mark.parametrize("shuffle", [False, True])
def test_rcm_permutation_colouring_tables(shuffle, order, cell_order):
    cell_order(order)
    m = M.Mesh.gen_tri(40, 30, 0, 4, 0, 3, 0.1, 3, (M.BC_INLET, M.BC_OUTLET, M.BC_PERIODIC, M.BC_PERIODIC))
    raw = m.raw()
    if shuffle:  # relabel cells randomly: RCM must recover a narrow band
        rng = np.random.default_rng(0)
        perm = rng.permutation(m.n_cells).astype(np.int32)
        raw["c0"] = perm[raw["c0"]]
        raw["c1"] = np.where(raw["c1"] >= 0, perm[np.maximum(raw["c1"], 0)], -1).astype(np.int32)
        m = M.Mesh.from_raw(raw)
    p = M.Prep(m)
    cn, fnw, col = p.permutation()
    assert np.array_equal(np.sort(cn), np.arange(m.n_cells))
    assert np.array_equal(np.sort(fnw), np.arange(m.n_faces))
    fc0, fc1, kind, ell = p.tables()
    # connectivity preserved bit-exactly under the permutation
    assert np.array_equal(fc0[fnw], cn[raw["c0"]])
    inter = raw["c1"] >= 0
    assert np.array_equal(fc1[fnw][inter], cn[raw["c1"][inter]])
    per = (raw["ftype"] == M.BC_PERIODIC) & (raw["partner"] > np.arange(m.n_faces))  # first face of a pair
    merged = (raw["ftype"] == M.BC_PERIODIC) & ~per
    assert np.array_equal(fc1[fnw][per], cn[raw["c0"][raw["partner"][per]]])
    assert (fc1[fnw][~inter & ~per] == -1).all()
    assert (kind[fnw][per] == 5).all() and (kind[fnw][merged] == 7).all()
    # cell->face slots list exactly the faces each cell accumulates: c0 side of every face but the
    # merged halves of periodic pairs, c1 side of interior faces and of periodic pairs
    inc = {}
    for f in range(m.n_faces):
        if kind[f] == 7:
            continue
        inc.setdefault(fc0[f], set()).add(2 * f)
        if kind[f] in (0, 5):
            inc.setdefault(fc1[f], set()).add(2 * f + 1)
    for c in range(m.n_cells):
        s = [e for e in ell[:, c] if e >= 0]
        assert sorted(s) == s and set(s) == inc[c]
        assert (ell[len(s):, c] == -1).all()
    # colouring: faces sharing an accumulating cell never share a colour
    seen = {}
    for f in range(m.n_faces):
        nf = fnw[f]
        if kind[nf] == 7:
            continue
        for c in ([fc0[nf], fc1[nf]] if kind[nf] in (0, 5) else [fc0[nf]]):
            key = (c, col[f])
            assert key not in seen
            seen[key] = f
    assert p.ncolours <= 5
    if order == M.ORDER_RCM:  # bandwidth of the new order vs the input order
        bw_new = np.abs(fc0[:p.n_interior] - fc1[:p.n_interior]).max()
        bw_old = np.abs(raw["c0"][inter] - raw["c1"][inter]).max()
        assert bw_new <= (bw_old if not shuffle else bw_old // 10)
    else:  # the order depends on the cell centres alone, not on the input labels
        assert block_internal_fraction(p) > 0.6


def block_internal_fraction(p, block=256):
    """Fraction of interior faces whose two cells lie in the same block of `block` consecutive cells."""
    fc0, fc1, kind, _ = p.tables()
    i = kind == 0
    return float(((fc0[i] // block) == (fc1[i] // block)).mean())


@pytest.mark.parametrize("n", [2, 4, 16])
def test_hilbert_order_is_a_unit_step_path_on_a_square_grid(n, cell_order):
    """On a 2^k x 2^k grid the Hilbert curve visits every cell once, each step to an edge neighbour,
    starting in the (0, 0) corner; RCM's diagonal fronts take longer steps."""
    cell_order(M.ORDER_HILBERT)
    m = M.Mesh.from_raw(quad_grid(n))
    cn, _, _ = M.Prep(m).permutation()
    old_of_new = np.argsort(cn)
    ij = np.array([(c % n, c // n) for c in old_of_new])
    step = np.abs(np.diff(ij, axis=0)).sum(1)
    assert (step == 1).all()
    assert tuple(ij[0]) == (0, 0)
    if n >= 4:
        cell_order(M.ORDER_RCM)
        cn, _, _ = M.Prep(m).permutation()
        ij = np.array([(c % n, c // n) for c in np.argsort(cn)])
        assert (np.abs(np.diff(ij, axis=0)).sum(1) > 1).any()


def test_hilbert_blocks_are_compact_in_3d(cell_order):
    """Tetrahedra: most interior faces join two cells of one 256-cell block along the Hilbert curve,
    few do along RCM's level sets."""
    m = M.Mesh.gen_tet(10, 10, 10, (0, 0, 0), (1, 1, 1), 0.1, 5, (M.BC_WALL,) * 6)
    cell_order(M.ORDER_HILBERT)
    hil = block_internal_fraction(M.Prep(m))
    cell_order(M.ORDER_RCM)
    rcm = block_internal_fraction(M.Prep(m))
    assert hil > 0.6 and hil > rcm + 0.2, (hil, rcm)


def test_cell_order_setting_rejects_unknown_values(cell_order):
    with pytest.raises(M.MstError):
        M.set_cell_order(7)
    assert M.get_cell_order() in (M.ORDER_HILBERT, M.ORDER_RCM)


def test_quad_cells_geometry_matches_oracle():
    fx = fixture("2d-stair-st-2")
    g = M.Prep(M.Mesh.from_raw(fx)).geometry()
    go = O.OracleMesh(fx).geometry()
    for k in g:
        assert np.array_equal(g[k], go[k], equal_nan=True), k
    # 20 sliver quads (width ~1e-8 along x = 0.6) get NaN from the reference's Heron split
    # (mesh/Cell.cpp:28-49); the device context refuses such a mesh instead of producing NaNs
    bad = ~(g["vol"] > 0)
    assert bad.sum() == 20 and np.isnan(g["vol"][bad]).all()


def test_brick_and_periodic_box_meshes():
    """The test meshes built from raw arrays: a running-bond brick wall (six-faced cells) and a tet box
    made periodic in x.  Geometry bit-identical to the oracle's; the cells tile the domain; every cell's
    outward normals close; the periodic partners pair faces of equal area."""
    for raw, area in ((brick_grid(10, 7, h=0.1, jitter=0.1, seed=2), 2.0 * 0.7),
                      (periodic_in_x(M.Mesh.gen_tet(5, 4, 3, (0, 0, 0), (1, 0.8, 0.7), 0.0, 3, (M.BC_WALL,) * 6).raw()),
                       1 * 0.8 * 0.7)):
        m = M.Mesh.from_raw(raw)
        g = M.Prep(m).geometry()
        go = O.OracleMesh(raw).geometry()
        for k in g:
            assert np.array_equal(g[k], go[k]), k
        assert abs(g["vol"].sum() - area) <= 1e-13 * area
        out = np.zeros((m.n_cells, int(raw["dim"])))
        sgn = g["dac"][:, None] * g["S"]  # the c0 outward normal
        np.add.at(out, raw["c0"], sgn)
        inner = raw["c1"] >= 0
        np.add.at(out, raw["c1"][inner], -sgn[inner])
        per = raw["ftype"] == 12
        # periodic faces close their c0 cell's surface with their own normal
        assert np.abs(out).max() <= 1e-14 * np.abs(g["S"]).max()
        if per.any():
            pa = raw["partner"][per]
            area_f = np.linalg.norm(g["S"], axis=1)
            assert np.allclose(area_f[per], area_f[pa], rtol=1e-14)
    nfaces = np.bincount(np.concatenate([raw["c0"], raw["c1"][raw["c1"] >= 0]]))
    assert nfaces.max() == 4  # (the tet box)
    b = brick_grid(10, 7)
    nb = np.bincount(np.concatenate([b["c0"], b["c1"][b["c1"] >= 0]]))
    assert sorted(set(nb.tolist())) == [4, 6]  # half bricks at the row ends, six-faced bricks
