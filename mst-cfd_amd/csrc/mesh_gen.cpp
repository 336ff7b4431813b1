// mesh_gen.cpp -- jittered simplex meshes for the synthetic workloads.
//
// The reference only reads Fluent files (mesh/MshBlock.cpp:5-18); these generators build the same
// file-order structure directly: interior faces first (one interior zone), then one zone per box
// side.  Every face's area vector points INTO its c0 cell, the orientation for which the reference's
// literal L/R rule (mesh/MshBlock.cpp:282-293) is the physically upwind one.
#include <algorithm>
#include <array>
#include <cmath>

#include "mesh.hpp"

namespace mst {

static bool valid_bc(int32_t t) {
    return t == MST_BC_WALL || t == MST_BC_OUTLET || t == MST_BC_SYMMETRY || t == MST_BC_INLET ||
           t == MST_BC_PERIODIC;
}

int gen_tri(int nx, int ny, double x0, double x1, double y0, double y1, double jitter, uint64_t seed,
            const int32_t bc[4], mst_mesh &m) {
    if (nx < 1 || ny < 1 || !(x1 > x0) || !(y1 > y0) || jitter < 0 || jitter >= 0.25)
        return fail(MST_E_ARG, "gen_tri: bad sizes or jitter (need nx,ny>=1, jitter in [0,0.25))");
    for (int s = 0; s < 4; s++)
        if (!valid_bc(bc[s])) return fail(MST_E_ARG, "gen_tri: unknown boundary type");
    bool perx = bc[0] == MST_BC_PERIODIC, pery = bc[2] == MST_BC_PERIODIC;
    if (perx != (bc[1] == MST_BC_PERIODIC) || pery != (bc[3] == MST_BC_PERIODIC))
        return fail(MST_E_ARG, "gen_tri: periodic sides must come in pairs");
    if ((int64_t)nx * ny * 2 > (int64_t)INT32_MAX / 4) return fail(MST_E_ARG, "gen_tri: mesh too large");
    m = mst_mesh();
    m.dim = 2;
    m.fnn = 2;
    const int64_t NX = nx + 1, NY = ny + 1;
    m.n_nodes = NX * NY;
    m.n_cells = 2LL * nx * ny;
    const double hx = (x1 - x0) / nx, hy = (y1 - y0) / ny;
    std::vector<double> jx(m.n_nodes), jy(m.n_nodes);
    Rng rng(seed);
    for (int64_t j = 0; j < NY; j++)
        for (int64_t i = 0; i < NX; i++) {
            double ax = rng.sym() * jitter * hx, ay = rng.sym() * jitter * hy;
            if (i == 0 || i == nx) ax = 0;  // boundary nodes move only along their side
            if (j == 0 || j == ny) ay = 0;
            jx[j * NX + i] = ax;
            jy[j * NX + i] = ay;
        }
    if (perx)
        for (int64_t j = 0; j < NY; j++) jy[j * NX + nx] = jy[j * NX + 0];
    if (pery)
        for (int64_t i = 0; i < NX; i++) jx[(int64_t)ny * NX + i] = jx[i];
    m.xyz.resize(m.n_nodes * 2);
    for (int64_t j = 0; j < NY; j++)
        for (int64_t i = 0; i < NX; i++) {
            m.xyz[(j * NX + i) * 2 + 0] = x0 + (double)i * hx + jx[j * NX + i];
            m.xyz[(j * NX + i) * 2 + 1] = y0 + (double)j * hy + jy[j * NX + i];
        }
    auto nid = [&](int64_t i, int64_t j) { return (int32_t)(j * NX + i); };
    auto tlo = [&](int64_t i, int64_t j) { return (int32_t)(2 * (j * nx + i)); };      // (a,b,c)
    auto thi = [&](int64_t i, int64_t j) { return (int32_t)(2 * (j * nx + i) + 1); };  // (a,c,d)
    auto add = [&](int32_t n0, int32_t n1, int32_t c0, int32_t c1, int32_t t) {
        m.fnodes.push_back(n0);
        m.fnodes.push_back(n1);
        m.c0.push_back(c0);
        m.c1.push_back(c1);
        m.ftype.push_back(t);
        m.partner.push_back(-1);
    };
    // interior: per quad the diagonal a->c, the right edge b->c, the top edge d->c
    for (int64_t j = 0; j < ny; j++)
        for (int64_t i = 0; i < nx; i++) {
            add(nid(i, j), nid(i + 1, j + 1), tlo(i, j), thi(i, j), MST_BC_INTERIOR);
            if (i + 1 < nx) add(nid(i + 1, j), nid(i + 1, j + 1), thi(i + 1, j), tlo(i, j), MST_BC_INTERIOR);
            if (j + 1 < ny) add(nid(i, j + 1), nid(i + 1, j + 1), thi(i, j), tlo(i, j + 1), MST_BC_INTERIOR);
        }
    m.n_int = (int64_t)m.c0.size();
    m.zones.push_back({3, 0, (int32_t)m.n_int, MST_BC_INTERIOR, "interior"});
    int64_t zx0 = m.c0.size();
    for (int64_t j = 0; j < ny; j++) add(nid(0, j), nid(0, j + 1), thi(0, j), -1, bc[0]);
    int64_t zx1 = m.c0.size();
    for (int64_t j = 0; j < ny; j++) add(nid(nx, j + 1), nid(nx, j), tlo(nx - 1, j), -1, bc[1]);
    int64_t zy0 = m.c0.size();
    for (int64_t i = 0; i < nx; i++) add(nid(i + 1, 0), nid(i, 0), tlo(i, 0), -1, bc[2]);
    int64_t zy1 = m.c0.size();
    for (int64_t i = 0; i < nx; i++) add(nid(i, ny), nid(i + 1, ny), thi(i, ny - 1), -1, bc[3]);
    int64_t zend = m.c0.size();
    m.zones.push_back({4, (int32_t)zx0, (int32_t)zx1, bc[0], "x0"});
    m.zones.push_back({5, (int32_t)zx1, (int32_t)zy0, bc[1], "x1"});
    m.zones.push_back({6, (int32_t)zy0, (int32_t)zy1, bc[2], "y0"});
    m.zones.push_back({7, (int32_t)zy1, (int32_t)zend, bc[3], "y1"});
    if (perx)
        for (int64_t j = 0; j < ny; j++) {
            m.partner[zx0 + j] = (int32_t)(zx1 + j);
            m.partner[zx1 + j] = (int32_t)(zx0 + j);
        }
    if (pery)
        for (int64_t i = 0; i < nx; i++) {
            m.partner[zy0 + i] = (int32_t)(zy1 + i);
            m.partner[zy1 + i] = (int32_t)(zy0 + i);
        }
    m.n_faces = (int64_t)m.c0.size();
    return validate(m);
}

// Compression ramp: the jittered triangle lattice of gen_tri on the unit square, mapped so that the
// bottom side is a wall with a corner at x = 0 (flat for x < 0, rising at theta for x > 0) and the top
// side is straight at y = height.  The column through the corner is pinned to it; the map keeps
// every cell's orientation, so faces still point into their c0 cell.
int gen_ramp(int nx, int ny, double x0, double x1, double height, double theta_deg, double jitter, uint64_t seed,
             const int32_t bc[4], mst_mesh &m) {
    if (!(x0 < 0) || !(x1 > 0) || !(theta_deg >= 0) || !(theta_deg < 45))
        return fail(MST_E_ARG, "gen_ramp: need x0 < 0 < x1 and 0 <= theta < 45 degrees");
    const double tn = std::tan(theta_deg * M_PI / 180.0);
    if (!(height > x1 * tn)) return fail(MST_E_ARG, "gen_ramp: the ramp reaches the top of the domain");
    const int ic = (int)std::lround(nx * (-x0) / (x1 - x0));
    if (ic < 1 || ic >= nx) return fail(MST_E_ARG, "gen_ramp: too few columns on one side of the corner");
    int rc = gen_tri(nx, ny, 0.0, 1.0, 0.0, 1.0, jitter, seed, bc, m);
    if (rc) return rc;
    const double xc = (double)ic / nx;
    m.xyz[(int64_t)ic * 2 + 0] = xc;  // node (ic, 0) is the corner
    m.xyz[(int64_t)ic * 2 + 1] = 0.0;
    for (int64_t n = 0; n < m.n_nodes; n++) {
        const double xi = m.xyz[n * 2], eta = m.xyz[n * 2 + 1];
        const double x = xi <= xc ? x0 * (1.0 - xi / xc) : x1 * (xi - xc) / (1.0 - xc);
        const double yb = x > 0 ? x * tn : 0.0;
        m.xyz[n * 2 + 0] = x;
        m.xyz[n * 2 + 1] = yb + (height - yb) * eta;
    }
    return validate(m);
}

// ---------------------------------------------------------------------------------------------- 3D
namespace {
struct FaceRec {
    int32_t a, b, c;  // sorted node ids
    int32_t slot;     // 4*tet + local face (the vertex left out)
    bool operator<(const FaceRec &o) const {
        if (a != o.a) return a < o.a;
        if (b != o.b) return b < o.b;
        if (c != o.c) return c < o.c;
        return slot < o.slot;
    }
    bool same(const FaceRec &o) const { return a == o.a && b == o.b && c == o.c; }
};
}  // namespace

int gen_tet(int nx, int ny, int nz, const double lo[3], const double hi[3], double jitter, uint64_t seed,
            const int32_t bc[6], mst_mesh &m) {
    if (nx < 1 || ny < 1 || nz < 1 || jitter < 0 || jitter >= 0.25)
        return fail(MST_E_ARG, "gen_tet: bad sizes or jitter");
    for (int s = 0; s < 6; s++) {
        if (!valid_bc(bc[s])) return fail(MST_E_ARG, "gen_tet: unknown boundary type");
        if (bc[s] == MST_BC_PERIODIC) return fail(MST_E_UNSUPPORTED, "gen_tet: periodic sides not supported");
    }
    if ((int64_t)nx * ny * nz * 6 > (int64_t)INT32_MAX / 4) return fail(MST_E_ARG, "gen_tet: mesh too large");
    m = mst_mesh();
    m.dim = 3;
    m.fnn = 3;
    const int64_t NX = nx + 1, NY = ny + 1, NZ = nz + 1;
    const int n[3] = {nx, ny, nz};
    m.n_nodes = NX * NY * NZ;
    const int64_t nhex = (int64_t)nx * ny * nz;
    m.n_cells = 6 * nhex;
    double h[3];
    for (int d = 0; d < 3; d++) h[d] = (hi[d] - lo[d]) / n[d];
    Rng rng(seed);
    m.xyz.resize(m.n_nodes * 3);
    for (int64_t k = 0; k < NZ; k++)
        for (int64_t j = 0; j < NY; j++)
            for (int64_t i = 0; i < NX; i++) {
                int64_t id = (k * NY + j) * NX + i;
                int64_t ijk[3] = {i, j, k};
                for (int d = 0; d < 3; d++) {
                    double a = rng.sym() * jitter * h[d];
                    if (ijk[d] == 0 || ijk[d] == n[d]) a = 0;
                    m.xyz[id * 3 + d] = lo[d] + (double)ijk[d] * h[d] + a;
                }
            }
    auto nid = [&](int64_t i, int64_t j, int64_t k) { return (int32_t)((k * NY + j) * NX + i); };
    // Kuhn split: tet p of hex (i,j,k) = path 0 -> e_a -> e_a+e_b -> 7 for permutation (a,b,c) p
    static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    std::vector<std::array<int32_t, 4>> tets(m.n_cells);
    for (int64_t k = 0; k < nz; k++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t i = 0; i < nx; i++) {
                int64_t hx = (k * ny + j) * nx + i;
                for (int p = 0; p < 6; p++) {
                    int64_t v[3] = {i, j, k};
                    std::array<int32_t, 4> t;
                    t[0] = nid(v[0], v[1], v[2]);
                    v[perm[p][0]]++;
                    t[1] = nid(v[0], v[1], v[2]);
                    v[perm[p][1]]++;
                    t[2] = nid(v[0], v[1], v[2]);
                    v[perm[p][2]]++;
                    t[3] = nid(v[0], v[1], v[2]);
                    tets[6 * hx + p] = t;
                }
            }
    std::vector<FaceRec> rec(4 * m.n_cells);
    for (int64_t t = 0; t < m.n_cells; t++)
        for (int l = 0; l < 4; l++) {
            int32_t q[3], w = 0;
            for (int s = 0; s < 4; s++)
                if (s != l) q[w++] = tets[t][s];
            std::sort(q, q + 3);
            rec[4 * t + l] = {q[0], q[1], q[2], (int32_t)(4 * t + l)};
        }
    std::sort(rec.begin(), rec.end());
    auto lattice = [&](int32_t node, double *x) {  // unjittered position (orientation decisions)
        int64_t i = node % NX, j = (node / NX) % NY, k = node / (NX * NY);
        x[0] = (double)i;
        x[1] = (double)j;
        x[2] = (double)k;
    };
    // S = 1/2 (p1-p0) x (p2-p0) must point into c0: order (a,b,c) so that S . (opposite - a) > 0
    auto oriented = [&](const FaceRec &r, int32_t tet, int32_t lface, int32_t out[3]) {
        double pa[3], pb[3], pc[3], po[3];
        lattice(r.a, pa);
        lattice(r.b, pb);
        lattice(r.c, pc);
        lattice(tets[tet][lface], po);
        double u[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
        double v[3] = {pc[0] - pa[0], pc[1] - pa[1], pc[2] - pa[2]};
        double cr[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        double s = cr[0] * (po[0] - pa[0]) + cr[1] * (po[1] - pa[1]) + cr[2] * (po[2] - pa[2]);
        out[0] = r.a;
        if (s > 0) {
            out[1] = r.b;
            out[2] = r.c;
        } else {
            out[1] = r.c;
            out[2] = r.b;
        }
    };
    std::vector<std::vector<FaceRec>> side(6);
    std::vector<std::pair<FaceRec, FaceRec>> inner;
    inner.reserve(2 * m.n_cells);
    for (size_t s = 0; s < rec.size();) {
        if (s + 1 < rec.size() && rec[s].same(rec[s + 1])) {
            inner.push_back({rec[s], rec[s + 1]});
            s += 2;
        } else {
            double pa[3], pb[3], pc[3];
            lattice(rec[s].a, pa);
            lattice(rec[s].b, pb);
            lattice(rec[s].c, pc);
            int which = -1;
            for (int d = 0; d < 3 && which < 0; d++) {
                if (pa[d] == 0 && pb[d] == 0 && pc[d] == 0) which = 2 * d;
                else if (pa[d] == n[d] && pb[d] == n[d] && pc[d] == n[d]) which = 2 * d + 1;
            }
            if (which < 0) return fail(MST_E_MESH, "gen_tet: unmatched interior face");
            side[which].push_back(rec[s]);
            s += 1;
        }
    }
    m.fnodes.reserve(3 * (inner.size() + 4 * (size_t)(nx * ny + ny * nz + nx * nz)));
    auto add = [&](const int32_t nd[3], int32_t c0, int32_t c1, int32_t t) {
        m.fnodes.push_back(nd[0]);
        m.fnodes.push_back(nd[1]);
        m.fnodes.push_back(nd[2]);
        m.c0.push_back(c0);
        m.c1.push_back(c1);
        m.ftype.push_back(t);
        m.partner.push_back(-1);
    };
    for (auto &pr : inner) {
        int32_t t0 = pr.first.slot / 4, t1 = pr.second.slot / 4;
        int32_t nd[3];
        oriented(pr.first, t0, pr.first.slot % 4, nd);
        add(nd, t0, t1, MST_BC_INTERIOR);
    }
    m.n_int = (int64_t)m.c0.size();
    m.zones.push_back({3, 0, (int32_t)m.n_int, MST_BC_INTERIOR, "interior"});
    static const char *names[6] = {"x0", "x1", "y0", "y1", "z0", "z1"};
    for (int s = 0; s < 6; s++) {
        int64_t first = m.c0.size();
        for (auto &r : side[s]) {
            int32_t t = r.slot / 4, nd[3];
            oriented(r, t, r.slot % 4, nd);
            add(nd, t, -1, bc[s]);
        }
        m.zones.push_back({4 + s, (int32_t)first, (int32_t)m.c0.size(), bc[s], names[s]});
    }
    m.n_faces = (int64_t)m.c0.size();
    return validate(m);
}

}  // namespace mst
