"""On-disk formats: the Tecplot result file (Work::writedataRhoBasedMshNodePlt, work/Work.cpp:204-304,
with its fixes) and the binary checkpoint (replacing the .cellPUVT dump, work/Work.cpp:432-514).  CPU only."""
import struct

import numpy as np
import pytest

import mstcfd as M
from helpers import GAMMA, cons, quad_grid

CV = 715.8  # include/CONST.h:39


def read_tecplot(path):
    lines = open(path).read().splitlines()
    assert lines[0].startswith('TITLE = "') and lines[0].endswith('"')
    names = [t.strip().strip('"') for t in lines[1].split("=", 1)[1].split(",")]
    zone = dict(kv.split("=") for kv in lines[2][len("ZONE "):].replace(" ", "").split(",") if "=" in kv)
    nn, ne = int(zone["NODES"]), int(zone["ELEMENTS"])
    data = np.array([[float(v) for v in l.split()] for l in lines[3:3 + nn]])
    conn = np.array([[int(v) for v in l.split()] for l in lines[3 + nn:3 + nn + ne]]) - 1
    assert len(lines) == 3 + nn + ne
    return names, zone, data, conn


def cell_node_sets(raw):
    sets = [set() for _ in range(int(raw["n_cells"]))]
    for f in range(len(raw["c0"])):
        for c in (raw["c0"][f], raw["c1"][f]):
            if c >= 0:
                sets[c].update(raw["fnodes"][f].tolist())
    return sets


def prim(Q, dim):
    rho = Q[:, 0]
    vel = Q[:, 1:1 + dim] / rho[:, None]
    e = Q[:, -1] - 0.5 * (Q[:, 1:1 + dim] ** 2).sum(1) / rho
    p = e * (GAMMA - 1)
    return rho, vel, e / rho / CV, p, np.sqrt((Q[:, 1:1 + dim] ** 2).sum(1) / (GAMMA * p * rho))


MESHES = {
    "tri": (lambda: M.Mesh.gen_tri(7, 5, 0, 2, 0, 1, 0.2, 3, (M.BC_INLET, M.BC_OUTLET, M.BC_WALL, M.BC_SYMMETRY)),
            "FETRIANGLE"),
    "periodic": (lambda: M.Mesh.gen_tri(6, 6, -1, 1, -1, 1, 0.1, 3, (M.BC_PERIODIC,) * 4), "FETRIANGLE"),
    "quad": (lambda: M.Mesh.from_raw(quad_grid(6, bc=M.BC_WALL, h=0.1, jitter=0.2, seed=4)), "FEQUADRILATERAL"),
    "tet": (lambda: M.Mesh.gen_tet(3, 2, 2, (0, 0, 0), (1, 0.8, 0.7), 0.15, 11, (M.BC_WALL,) * 6), "FETETRAHEDRON"),
}


@pytest.mark.parametrize("name", sorted(MESHES))
def test_tecplot_layout_and_uniform_state(name, tmp_path):
    mk, zt = MESHES[name]
    m = mk()
    raw = m.raw()
    dim = m.dim
    q = cons(np.array([1.3]), np.array([[0.4, -0.2, 0.1][:dim]]), np.array([0.9]))[0]
    Q = np.tile(q, (m.n_cells, 1))
    path = tmp_path / "out.plt"
    m.write_tecplot(Q, path, title="uniform")
    names, zone, data, conn = read_tecplot(path)
    assert names == ["X", "Y", "Z"][:dim] + ["rho"] + ["u", "v", "w"][:dim] + ["T", "p", "Ma"]
    assert zone["ZONETYPE"] == zt and int(zone["NODES"]) == m.n_nodes and int(zone["ELEMENTS"]) == m.n_cells
    assert data.shape == (m.n_nodes, 2 * dim + 4)
    assert np.array_equal(data[:, :dim], raw["xyz"])  # coordinates round-trip exactly (17 digits)
    rho, vel, T, p, ma = prim(q[None, :], dim)
    expect = np.concatenate([rho, vel[0], T, p, ma])
    assert np.allclose(data[:, dim:], expect[None, :], rtol=1e-14, atol=0)
    # element node lists are the cells' nodes; quadrilaterals in cyclic order
    sets = cell_node_sets(raw)
    for c in range(m.n_cells):
        assert set(conn[c].tolist()) == sets[c]
    if zt == "FEQUADRILATERAL":
        edges = {frozenset(raw["fnodes"][f].tolist()) for f in range(len(raw["c0"]))}
        for c in range(m.n_cells):
            for k in range(4):
                assert frozenset((conn[c][k], conn[c][(k + 1) % 4])) in edges


def test_tecplot_node_values_are_face_area_weighted():
    """Node value = sum over the node's faces of w_f q_f / sum w_f, w_f = 1/|S_f| of that face (the
    reference weights by the face whose index equals the node's, work/Work.cpp:292-293); q_f = the
    eta0-weighted mean of the face's cells (boundary: its cell).  Stretched cells make the two
    weightings differ."""
    m = M.Mesh.gen_tri(5, 4, 0, 4, 0, 0.5, 0.2, 7, (M.BC_INLET, M.BC_OUTLET, M.BC_WALL, M.BC_SYMMETRY))
    raw = m.raw()
    geo = M.Prep(m).geometry()
    rng = np.random.default_rng(1)
    Q = cons(rng.uniform(0.5, 2, m.n_cells), rng.uniform(-0.5, 0.5, (m.n_cells, 2)), rng.uniform(0.5, 2, m.n_cells))
    import tempfile, os
    with tempfile.TemporaryDirectory() as d:
        m.write_tecplot(Q, os.path.join(d, "a.plt"))
        _, _, data, _ = read_tecplot(os.path.join(d, "a.plt"))
    nf = len(raw["c0"])
    qf = np.empty((nf, 4))
    for f in range(nf):
        c0, c1 = raw["c0"][f], raw["c1"][f]
        qf[f] = geo["eta"][f] * Q[c0] + (1 - geo["eta"][f]) * Q[c1] if c1 >= 0 else Q[c0]
    w = 1 / np.linalg.norm(geo["S"], axis=1)
    acc = np.zeros((m.n_nodes, 4))
    ws = np.zeros(m.n_nodes)
    for f in range(nf):
        for v in raw["fnodes"][f]:
            acc[v] += w[f] * qf[f]
            ws[v] += w[f]
    qn = acc / ws[:, None]
    rho, vel, T, p, ma = prim(qn, 2)
    assert np.allclose(data[:, 2], rho, rtol=1e-13)
    assert np.allclose(data[:, 3:5], vel, rtol=1e-12, atol=1e-14)
    assert np.allclose(data[:, 5:], np.stack([T, p, ma], 1), rtol=1e-12)
    # each node value is a convex combination of the adjacent cells' densities
    sets = cell_node_sets(raw)
    lo = np.full(m.n_nodes, np.inf)
    hi = np.full(m.n_nodes, -np.inf)
    for c, s in enumerate(sets):
        for v in s:
            lo[v] = min(lo[v], Q[c, 0])
            hi[v] = max(hi[v], Q[c, 0])
    assert np.all(data[:, 2] >= lo - 1e-12) and np.all(data[:, 2] <= hi + 1e-12)


def test_tecplot_wall_noslip_zeroes_wall_face_velocity(tmp_path):
    m = M.Mesh.gen_tri(4, 3, 0, 1, 0, 1, 0.0, 1, (M.BC_WALL,) * 4)
    Q = np.tile(cons(np.array([1.0]), np.array([[0.5, 0.2]]), np.array([1.0]))[0], (m.n_cells, 1))
    m.write_tecplot(Q, tmp_path / "a.plt", wall_noslip=False)
    m.write_tecplot(Q, tmp_path / "b.plt", wall_noslip=True)
    _, _, a, _ = read_tecplot(tmp_path / "a.plt")
    _, _, b, _ = read_tecplot(tmp_path / "b.plt")
    assert np.allclose(a[:, 3:5], [0.5, 0.2], rtol=1e-14)
    raw = m.raw()
    on_wall = np.zeros(m.n_nodes, bool)
    on_wall[raw["fnodes"][raw["c1"] < 0].ravel()] = True
    # nodes on a wall face average in its zero velocity; the others keep the cells' velocity
    assert np.all(np.abs(b[on_wall, 3:5]) < np.abs(a[on_wall, 3:5]))
    assert np.allclose(b[~on_wall, 3:5], [0.5, 0.2], rtol=1e-14)
    assert np.allclose(b[:, 2], a[:, 2], rtol=1e-14)  # density unchanged


def test_checkpoint_round_trip_and_refusals(tmp_path):
    m = M.Mesh.gen_tri(9, 7, 0, 1, 0, 1, 0.1, 5, (M.BC_WALL,) * 4)
    rng = np.random.default_rng(3)
    Q = rng.standard_normal((m.n_cells, 4))
    path = tmp_path / "run.ckpt"
    m.write_checkpoint(path, Q, step=1234, time=0.125)
    assert not (tmp_path / "run.ckpt.tmp").exists()
    Q2, step, t = m.read_checkpoint(path)
    assert np.array_equal(Q2, Q) and step == 1234 and t == 0.125
    raw = path.read_bytes()
    # a flipped bit in the state
    bad = bytearray(raw)
    bad[-9] ^= 0x10
    (tmp_path / "bad.ckpt").write_bytes(bytes(bad))
    with pytest.raises(M.MstError, match="checksum"):
        m.read_checkpoint(tmp_path / "bad.ckpt")
    # truncated
    (tmp_path / "short.ckpt").write_bytes(raw[:-8])
    with pytest.raises(M.MstError, match="truncated"):
        m.read_checkpoint(tmp_path / "short.ckpt")
    # another mesh of the same size (different jitter)
    other = M.Mesh.gen_tri(9, 7, 0, 1, 0, 1, 0.1, 6, (M.BC_WALL,) * 4)
    with pytest.raises(M.MstError, match="another mesh"):
        other.read_checkpoint(path)
    with pytest.raises(M.MstError):
        m.read_checkpoint(tmp_path / "missing.ckpt")
    (tmp_path / "text.ckpt").write_bytes(b"not a checkpoint at all" * 10)
    with pytest.raises(M.MstError, match="format"):
        m.read_checkpoint(tmp_path / "text.ckpt")
    # a header edited to 5 fields per cell, with 5 fields of data and their checksum: only the field
    # count can refuse it (the caller's buffer holds dim + 2 per cell)
    def fnv1a(b):
        h = 1469598103934665603
        for x in b:
            h = ((h ^ x) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
        return h
    head = bytearray(raw[:80])
    assert struct.unpack_from("<I", head, 16)[0] == 4  # nv
    struct.pack_into("<I", head, 16, 5)
    data = rng.standard_normal(m.n_cells * 5).tobytes()
    struct.pack_into("<Q", head, 72, fnv1a(data))
    (tmp_path / "nv5.ckpt").write_bytes(bytes(head) + data)
    with pytest.raises(M.MstError, match="format"):
        m.read_checkpoint(tmp_path / "nv5.ckpt")
    # the same edit back to 4 fields reads (the forging is right)
    struct.pack_into("<I", head, 16, 4)
    data4 = data[: m.n_cells * 4 * 8]
    struct.pack_into("<Q", head, 72, fnv1a(data4))
    (tmp_path / "nv4.ckpt").write_bytes(bytes(head) + data4)
    assert np.array_equal(m.read_checkpoint(tmp_path / "nv4.ckpt")[0].reshape(-1),
                          np.frombuffer(data4, np.float64))


@pytest.mark.parametrize("name", ["tri", "periodic", "quad", "tet", "fluent"])
def test_mesh_cache_round_trip(name, tmp_path):
    if name == "fluent":
        import os
        d = dict(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "msh_2d-stairW-1.npz")))
        for k in ("dim", "n_cells", "n_int"):
            d[k] = int(d[k])
        m = M.Mesh.from_raw(d)
    else:
        m = MESHES[name][0]()
    path = tmp_path / "mesh.bin"
    m.write_cache(path)
    m2 = M.Mesh.read_cache(path)
    a, b = m.raw(), m2.raw()
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    assert np.array_equal(m.zones(), m2.zones())
    assert (m2.n_cells, m2.n_faces, m2.n_nodes, m2.n_int) == (m.n_cells, m.n_faces, m.n_nodes, m.n_int)
    # a damaged copy, a truncated copy and another format are refused
    raw = path.read_bytes()
    bad = bytearray(raw)
    bad[len(raw) // 2] ^= 0x01
    (tmp_path / "bad.bin").write_bytes(bytes(bad))
    with pytest.raises(M.MstError, match="checksum"):
        M.Mesh.read_cache(tmp_path / "bad.bin")
    (tmp_path / "short.bin").write_bytes(raw[:len(raw) // 3])
    with pytest.raises(M.MstError):
        M.Mesh.read_cache(tmp_path / "short.bin")
    m.write_checkpoint(tmp_path / "state.ckpt", np.ones((m.n_cells, m.dim + 2)))
    with pytest.raises(M.MstError, match="format"):
        M.Mesh.read_cache(tmp_path / "state.ckpt")
