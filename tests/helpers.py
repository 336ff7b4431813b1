"""Shared inputs and parity checks for the tests (states, meshes, tolerance checks)."""
import numpy as np

GAMMA = 1.4


def cons(rho, vel, p, gamma=GAMMA):
    """Conserved state from primitives; vel has shape (n, dim)."""
    rho = np.asarray(rho, np.float64)
    vel = np.asarray(vel, np.float64)
    p = np.asarray(p, np.float64)
    n, dim = vel.shape
    Q = np.zeros((n, dim + 2))
    Q[:, 0] = rho
    Q[:, 1:dim + 1] = rho[:, None] * vel
    Q[:, dim + 1] = p / (gamma - 1) + 0.5 * rho * (vel ** 2).sum(1)
    return Q


def prim(Q, gamma=GAMMA):
    dim = Q.shape[1] - 2
    rho = Q[:, 0]
    vel = Q[:, 1:dim + 1] / rho[:, None]
    p = (gamma - 1) * (Q[:, dim + 1] - 0.5 * rho * (vel ** 2).sum(1))
    return rho, vel, p


def sod_state(ccen, x_split=0.5, left=(1.0, 1.0), right=(0.125, 0.1), gamma=GAMMA):
    n, dim = ccen.shape
    L = ccen[:, 0] <= x_split
    rho = np.where(L, left[0], right[0])
    p = np.where(L, left[1], right[1])
    return cons(rho, np.zeros((n, dim)), p, gamma)


def smooth_state(ccen, seed=0, amp=0.3, mach=0.8, gamma=GAMMA):
    """Smooth, everywhere-physical field with sub- and supersonic regions (mach is the peak speed/a)."""
    rng = np.random.default_rng(seed)
    n, dim = ccen.shape
    k = rng.uniform(2, 7, size=(4, dim))
    ph = rng.uniform(0, 2 * np.pi, size=4)
    x = ccen / (np.abs(ccen).max(0) + 1e-30)
    w = [np.sin(x @ k[i] * np.pi + ph[i]) for i in range(4)]
    rho = 1 + amp * w[0]
    p = 1 + amp * w[1]
    a = np.sqrt(gamma * p / rho)
    vel = np.zeros((n, dim))
    vel[:, 0] = mach * a * w[2]
    vel[:, 1] = mach * a * w[3]
    if dim == 3:
        vel[:, 2] = 0.5 * mach * a * w[2] * w[3]
    return cons(rho, vel, p, gamma)


def jumpy_state(n, dim=2, seed=0, gamma=GAMMA, vmax=2.0):
    """Independent random states per cell: large jumps, sub/supersonic, both flow directions."""
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.3, 3.0, n)
    p = rng.uniform(0.2, 3.0, n)
    a = np.sqrt(gamma * p / rho)
    vel = rng.uniform(-vmax, vmax, (n, dim)) * a[:, None]
    return cons(rho, vel, p, gamma)


def flux_scale(Q, gamma=GAMMA):
    """Natural flux magnitude of each cell's state per field: rho c (1, c.., H), c = |u| + a."""
    rho, vel, p = prim(Q, gamma)
    dim = vel.shape[1]
    c = np.sqrt((vel ** 2).sum(1)) + np.sqrt(gamma * np.abs(p) / rho)
    H = (Q[:, dim + 1] + p) / rho
    s = np.empty_like(Q)
    s[:, 0] = rho * c
    s[:, 1:dim + 1] = (rho * c * c)[:, None]
    s[:, dim + 1] = rho * c * np.abs(H)
    return s


def normal_sums(raw, S):
    """sum over a cell's faces of sum_d |S_f,d| (file order)."""
    w = np.abs(S).sum(1)
    out = np.zeros(int(raw["n_cells"]))
    np.add.at(out, raw["c0"], w)
    m = raw["c1"] >= 0
    np.add.at(out, raw["c1"][m], w[m])
    return out


def residual_parity_error(Rg, Rr, Sr, floor=None):
    """max over cells and fields of |R_gpu - R_ref| / S_ref, with S_ref = sum_f sum_d |n_d| |F_d|
    (the parity scale).

    floor (flux_scale(Q) * normal_sums: the natural flux magnitude of the cell's state through its
    faces) is added to S_ref only in the cells and fields where S_ref is itself round-off, below 1e-6
    of it: a field whose every face flux vanishes (the mass and energy fluxes of gas at rest) has no
    scale of its own.  Everywhere else the bar is relative to S_ref alone."""
    den = Sr
    if floor is not None:
        den = np.where(Sr <= 1e-6 * floor, Sr + floor, Sr)
    den = np.where(den > 0, den, 1.0)
    err = np.abs(Rg - Rr) / den
    return float(err.max())


def rel_state_error(Qg, Qr, gamma=GAMMA):
    """max over cells of |Q_gpu - Q_ref| / field scale; scales: max rho, max rho(|u|+a) for every
    momentum component, max rho E."""
    rho, vel, p = prim(Qr, gamma)
    dim = vel.shape[1]
    c = np.sqrt((vel ** 2).sum(1)) + np.sqrt(gamma * np.abs(p) / rho)
    sc = np.empty(dim + 2)
    sc[0] = np.abs(rho).max()
    sc[1:dim + 1] = (rho * c).max()
    sc[dim + 1] = np.abs(Qr[:, dim + 1]).max()
    return float((np.abs(Qg - Qr) / sc[None, :]).max())


def flip_faces(raw, frac=0.5, seed=0):
    """Reverse the node order of a random subset of 2D faces: their area vector then points out of c0
    (the reference reader's "branch 2", as in all bundled Fluent meshes)."""
    rng = np.random.default_rng(seed)
    out = dict(raw)
    fn = raw["fnodes"].copy()
    sel = rng.random(fn.shape[0]) < frac
    fn[sel] = fn[sel][:, ::-1]
    out["fnodes"] = np.ascontiguousarray(fn)
    return out, sel


def quad_grid(n, bc=3, h=1.0, jitter=0.0, seed=0):
    """n x n quadrilaterals of size h (interior nodes moved by up to jitter * h), interior faces
    first, boundary faces of Fluent type bc around; the raw arrays of Mesh.from_raw."""
    nid = lambda i, j: j * (n + 1) + i
    cid = lambda i, j: j * n + i
    xy = np.array([(i, j) for j in range(n + 1) for i in range(n + 1)], np.float64) * h
    if jitter:
        rng = np.random.default_rng(seed)
        inner = np.array([0 < i < n and 0 < j < n for j in range(n + 1) for i in range(n + 1)])
        xy[inner] += rng.uniform(-jitter * h, jitter * h, (int(inner.sum()), 2))
    fn, c0, c1 = [], [], []
    for j in range(n):
        for i in range(n - 1):
            fn.append((nid(i + 1, j), nid(i + 1, j + 1)))
            c0.append(cid(i, j))
            c1.append(cid(i + 1, j))
    for j in range(n - 1):
        for i in range(n):
            fn.append((nid(i, j + 1), nid(i + 1, j + 1)))
            c0.append(cid(i, j))
            c1.append(cid(i, j + 1))
    nint = len(fn)
    for k in range(n):
        for a, b, c in ((nid(k, 0), nid(k + 1, 0), cid(k, 0)), (nid(k, n), nid(k + 1, n), cid(k, n - 1)),
                        (nid(0, k), nid(0, k + 1), cid(0, k)), (nid(n, k), nid(n, k + 1), cid(n - 1, k))):
            fn.append((a, b))
            c0.append(c)
            c1.append(-1)
    nf = len(fn)
    return dict(dim=2, xyz=xy, fnodes=np.array(fn, np.int32), c0=np.array(c0, np.int32), c1=np.array(c1, np.int32),
                ftype=np.array([2] * nint + [bc] * (nf - nint), np.int32), partner=np.full(nf, -1, np.int32),
                n_cells=n * n, n_int=nint)


def periodic_in_x(raw):
    """The raw arrays of a mesh whose boundary faces on its lowest and highest x planes match in
    (y, z) (an unjittered gen_tet box), with those faces paired as periodic partners (Fluent type
    12): the 3D periodic meshes the generators do not make."""
    out = {k: np.array(v) for k, v in raw.items()}
    cen = out["xyz"][out["fnodes"]].mean(1)
    bnd = out["c1"] < 0
    x0, x1 = out["xyz"][:, 0].min(), out["xyz"][:, 0].max()
    lo = np.where(bnd & (np.abs(cen[:, 0] - x0) < 1e-12))[0]
    hi = np.where(bnd & (np.abs(cen[:, 0] - x1) < 1e-12))[0]
    key = lambda i: tuple(np.round(cen[i, 1:], 9))
    pair = {key(j): j for j in hi}
    assert len(lo) == len(hi) == len(pair) and len(lo) > 0
    for i in lo:
        j = pair[key(i)]
        out["partner"][i], out["partner"][j] = j, i
        out["ftype"][i] = out["ftype"][j] = 12
    return out


def brick_grid(m, n, h=1.0, jitter=0.0, seed=0, bc=3):
    """A running-bond brick wall: n rows of bricks 2h wide and h high, odd rows shifted by h (a half
    brick at each end), so a full brick has six faces (two below, two above, one each side) and the
    cell tables need eight slots.  Interior nodes move by up to jitter * h; boundary faces of Fluent
    type bc.  The raw arrays of Mesh.from_raw, interior faces first."""
    nid = lambda i, j: j * (2 * m + 1) + i
    xy = np.array([(i, j) for j in range(n + 1) for i in range(2 * m + 1)], np.float64) * h
    if jitter:
        rng = np.random.default_rng(seed)
        inner = np.array([0 < i < 2 * m and 0 < j < n for j in range(n + 1) for i in range(2 * m + 1)])
        xy[inner] += rng.uniform(-jitter * h, jitter * h, (int(inner.sum()), 2))
    first, count = [], 0  # the first cell of each row
    for j in range(n):
        first.append(count)
        count += m if j % 2 == 0 else m + 1
    cell = lambda x, j: first[j] + (x // 2 if j % 2 == 0 else (x + 1) // 2)  # the row-j cell over [x, x+1]
    fn, c0, c1 = [], [], []
    for j in range(n):  # vertical faces inside a row
        for x in range(1, 2 * m):
            if (x % 2 == 0) == (j % 2 == 0):
                fn.append((nid(x, j), nid(x, j + 1)))
                c0.append(cell(x - 1, j))
                c1.append(cell(x, j))
    for j in range(1, n):  # unit segments between rows j - 1 and j
        for x in range(2 * m):
            fn.append((nid(x, j), nid(x + 1, j)))
            c0.append(cell(x, j - 1))
            c1.append(cell(x, j))
    nint = len(fn)
    for x in range(2 * m):
        fn.append((nid(x, 0), nid(x + 1, 0)))
        c0.append(cell(x, 0))
        c1.append(-1)
        fn.append((nid(x, n), nid(x + 1, n)))
        c0.append(cell(x, n - 1))
        c1.append(-1)
    for j in range(n):
        fn.append((nid(0, j), nid(0, j + 1)))
        c0.append(cell(0, j))
        c1.append(-1)
        fn.append((nid(2 * m, j), nid(2 * m, j + 1)))
        c0.append(cell(2 * m - 1, j))
        c1.append(-1)
    nf = len(fn)
    return dict(dim=2, xyz=xy, fnodes=np.array(fn, np.int32), c0=np.array(c0, np.int32), c1=np.array(c1, np.int32),
                ftype=np.array([2] * nint + [bc] * (nf - nint), np.int32), partner=np.full(nf, -1, np.int32),
                n_cells=count, n_int=nint)
