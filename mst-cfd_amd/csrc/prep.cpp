// prep.cpp -- geometry, Hilbert-curve or reverse Cuthill-McKee cell order, face order and colouring,
// SoA arrays.
//
// Geometry follows the reference exactly (same operations in the same order, FP64, built with
// -ffp-contract=off so it is bit-identical wherever it is compiled):
//   face centre = node mean, 2D area vector S = (-(y0-y1), x0-x1)        mesh/Face.cpp:8-19
//   3D triangle S = 1/2 (p1-p0) x (p2-p0)                                mesh/Face.cpp:23-28
//   cell centre = mean of its face centres                               mesh/Cell.cpp:6-14
//   triangle area by Heron, quad by the "middle line" split               mesh/Cell.cpp:15-49
//   directAndCells = -1 when S . (x_f - x_c0) < 0                        mesh/MshBlock.cpp:281-283
//   outward normal of c0 = directAndCells * S, of c1 = -that              mesh/MshBlock.cpp:307-318
//   eta0 = |x_c1 - x_f| / (|x_c0 - x_f| + |x_c1 - x_f|), 1 on boundaries  mesh/Face.cpp:62-69
// New here: volumes of other polytopes by the divergence theorem (the reference's 3D formula,
// mesh/Cell.cpp:52-59, lacks the 1/3 and is orientation dependent), eta0 of periodic faces through
// the partner face, Hilbert and RCM orderings, face colouring and the slot-major cell->face table.
#include "prep.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <numeric>

namespace mst {

namespace {

inline double dotn(const double *a, const double *b, int dim) {
    double s = a[0] * b[0] + a[1] * b[1];
    if (dim == 3) s += a[2] * b[2];
    return s;
}
inline double normn(const double *a, int dim) { return std::sqrt(dotn(a, a, dim)); }

void geometry(const mst_mesh &m, Prepared &P, std::vector<int64_t> &cptr, std::vector<int32_t> &cface) {
    const int dim = m.dim;
    const int64_t nf = m.n_faces, nc = m.n_cells;
    // cell -> faces in reader attachment order (MshBlock.cpp:238-255)
    cptr.assign(nc + 1, 0);
    for (int64_t f = 0; f < nf; f++) {
        cptr[m.c0[f] + 1]++;
        if (m.c1[f] >= 0) cptr[m.c1[f] + 1]++;
    }
    for (int64_t c = 0; c < nc; c++) cptr[c + 1] += cptr[c];
    cface.assign(cptr[nc], 0);
    {
        std::vector<int64_t> fill(cptr.begin(), cptr.end() - 1);
        for (int64_t f = 0; f < nf; f++) {
            cface[fill[m.c0[f]]++] = (int32_t)f;
            if (m.c1[f] >= 0) cface[fill[m.c1[f]]++] = (int32_t)f;
        }
    }
    P.fcen.assign(nf * dim, 0);
    P.S.assign(nf * dim, 0);
    for (int64_t f = 0; f < nf; f++) {
        double sum[3] = {0, 0, 0};
        for (int j = 0; j < m.fnn; j++) {
            const double *p = &m.xyz[(int64_t)m.fnodes[f * m.fnn + j] * dim];
            for (int k = 0; k < dim; k++) sum[k] += p[k];
        }
        for (int k = 0; k < dim; k++) P.fcen[f * dim + k] = sum[k] / (double)m.fnn;
        const double *p0 = &m.xyz[(int64_t)m.fnodes[f * m.fnn + 0] * dim];
        const double *p1 = &m.xyz[(int64_t)m.fnodes[f * m.fnn + 1] * dim];
        if (dim == 2) {
            double fx = p0[0] - p1[0], fy = p0[1] - p1[1];
            P.S[f * 2 + 0] = -fy;
            P.S[f * 2 + 1] = fx;
        } else {
            const double *p2 = &m.xyz[(int64_t)m.fnodes[f * m.fnn + 2] * dim];
            double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
            double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
            double cr[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
            double s = m.fnn == 3 ? 0.5 : 1.0;
            for (int k = 0; k < 3; k++) P.S[f * 3 + k] = s * cr[k];
        }
    }
    P.ccen.assign(nc * dim, 0);
    for (int64_t c = 0; c < nc; c++) {
        double sum[3] = {0, 0, 0};
        for (int64_t j = cptr[c]; j < cptr[c + 1]; j++)
            for (int k = 0; k < dim; k++) sum[k] += P.fcen[(int64_t)cface[j] * dim + k];
        for (int k = 0; k < dim; k++) P.ccen[c * dim + k] = sum[k] / (double)(cptr[c + 1] - cptr[c]);
    }
    P.dac.assign(nf, 1);
    for (int64_t f = 0; f < nf; f++) {
        double dx[3] = {0, 0, 0};
        for (int k = 0; k < dim; k++) dx[k] = P.fcen[f * dim + k] - P.ccen[(int64_t)m.c0[f] * dim + k];
        P.dac[f] = dotn(&P.S[f * dim], dx, dim) < 0 ? -1 : 1;
    }
    P.vol.assign(nc, 0);
    P.cshalf.assign(nc * dim, 0);  // file order for now, permuted below
    for (int64_t c = 0; c < nc; c++) {
        const int64_t b = cptr[c], k = cptr[c + 1] - b;
        double V = 0;
        if (dim == 2 && k == 3) {
            double l[3];
            for (int i = 0; i < 3; i++) l[i] = normn(&P.S[(int64_t)cface[b + i] * 2], 2);
            double s = 0.5 * (l[0] + l[1] + l[2]);
            V = std::sqrt(s * (s - l[0]) * (s - l[1]) * (s - l[2]));
        } else if (dim == 2 && k == 4) {
            const double *fc[4];
            double L[4];
            for (int i = 0; i < 4; i++) {
                fc[i] = &P.fcen[(int64_t)cface[b + i] * 2];
                L[i] = normn(&P.S[(int64_t)cface[b + i] * 2], 2);
            }
            double v1[2] = {fc[0][0] - fc[1][0], fc[0][1] - fc[1][1]};
            double v2[2] = {fc[2][0] - fc[3][0], fc[2][1] - fc[3][1]};
            int i0, i1, i2, i3;
            double ml;
            if (std::fabs(v1[0] * v2[1] - v2[0] * v1[1]) < 1e-7) {
                ml = 2 * normn(v1, 2);
                i0 = 0; i1 = 1; i2 = 2; i3 = 3;
            } else {
                double v3[2] = {fc[0][0] - fc[2][0], fc[0][1] - fc[2][1]};
                ml = 2 * normn(v3, 2);
                i0 = 0; i1 = 2; i2 = 1; i3 = 3;
            }
            double s = 0.5 * (L[i0] + L[i1] + ml);
            V = std::sqrt(s * (s - L[i0]) * (s - L[i1]) * (s - ml));
            double s2 = 0.5 * (L[i2] + L[i3] + ml);
            V += std::sqrt(s2 * (s2 - L[i2]) * (s2 - L[i3]) * (s2 - ml));
        } else {
            for (int64_t j = b; j < b + k; j++) {
                int32_t f = cface[j];
                double sg = (c == m.c0[f]) ? (double)P.dac[f] : -(double)P.dac[f];
                double dx[3] = {0, 0, 0};
                for (int q = 0; q < dim; q++) dx[q] = P.fcen[(int64_t)f * dim + q] - P.ccen[c * dim + q];
                V += sg * dotn(&P.S[(int64_t)f * dim], dx, dim);
            }
            V /= dim;
        }
        P.vol[c] = V;
        for (int64_t j = b; j < b + k; j++)
            for (int q = 0; q < dim; q++) P.cshalf[c * dim + q] += std::fabs(P.S[(int64_t)cface[j] * dim + q]);
        for (int q = 0; q < dim; q++) P.cshalf[c * dim + q] *= 0.5;
    }
    P.eta.assign(nf, 1.0);
    for (int64_t f = 0; f < nf; f++) {
        const double *xf = &P.fcen[f * dim];
        double d0[3] = {0, 0, 0}, d1[3] = {0, 0, 0};
        if (m.c1[f] >= 0) {
            for (int k = 0; k < dim; k++) {
                d0[k] = P.ccen[(int64_t)m.c0[f] * dim + k] - xf[k];
                d1[k] = P.ccen[(int64_t)m.c1[f] * dim + k] - xf[k];
            }
            P.eta[f] = normn(d1, dim) / (normn(d0, dim) + normn(d1, dim));
        } else if (m.ftype[f] == MST_BC_PERIODIC && m.partner[f] >= 0) {
            int32_t g = m.partner[f];
            for (int k = 0; k < dim; k++) {
                d0[k] = P.ccen[(int64_t)m.c0[f] * dim + k] - xf[k];
                d1[k] = P.ccen[(int64_t)m.c0[g] * dim + k] - P.fcen[(int64_t)g * dim + k];
            }
            P.eta[f] = normn(d1, dim) / (normn(d0, dim) + normn(d1, dim));
        }
    }
}

// Reverse Cuthill-McKee over the cell graph (interior faces and periodic pairs), George-Liu
// pseudo-peripheral start per connected component; neighbours visited by (degree, id).
void rcm(const mst_mesh &m, std::vector<int32_t> &old_of_new) {
    const int64_t nc = m.n_cells;
    std::vector<int64_t> ap(nc + 1, 0);
    auto nbr_of = [&](int64_t f, int32_t &a, int32_t &b) -> bool {
        if (m.c1[f] >= 0) {
            a = m.c0[f];
            b = m.c1[f];
            return true;
        }
        if (m.ftype[f] == MST_BC_PERIODIC && m.partner[f] > f) {
            a = m.c0[f];
            b = m.c0[m.partner[f]];
            return a != b;
        }
        return false;
    };
    for (int64_t f = 0; f < m.n_faces; f++) {
        int32_t a, b;
        if (nbr_of(f, a, b)) {
            ap[a + 1]++;
            ap[b + 1]++;
        }
    }
    for (int64_t c = 0; c < nc; c++) ap[c + 1] += ap[c];
    std::vector<int32_t> adj(ap[nc]);
    {
        std::vector<int64_t> fill(ap.begin(), ap.end() - 1);
        for (int64_t f = 0; f < m.n_faces; f++) {
            int32_t a, b;
            if (nbr_of(f, a, b)) {
                adj[fill[a]++] = b;
                adj[fill[b]++] = a;
            }
        }
    }
    std::vector<int32_t> deg(nc);
    for (int64_t c = 0; c < nc; c++) deg[c] = (int32_t)(ap[c + 1] - ap[c]);
    for (int64_t c = 0; c < nc; c++)
        std::sort(adj.begin() + ap[c], adj.begin() + ap[c + 1], [&](int32_t x, int32_t y) {
            return deg[x] != deg[y] ? deg[x] < deg[y] : x < y;
        });
    std::vector<int32_t> level(nc, -1);
    std::vector<char> done(nc, 0);
    std::vector<int32_t> queue;
    queue.reserve(nc);
    // BFS from s over not-yet-numbered cells; returns the last level's min-degree cell and the depth
    auto bfs_probe = [&](int32_t s, int32_t &far, int32_t &depth) {
        std::vector<int32_t> touched;
        std::vector<int32_t> q{s};
        level[s] = 0;
        touched.push_back(s);
        size_t h = 0;
        while (h < q.size()) {
            int32_t u = q[h++];
            for (int64_t j = ap[u]; j < ap[u + 1]; j++) {
                int32_t v = adj[j];
                if (!done[v] && level[v] < 0) {
                    level[v] = level[u] + 1;
                    q.push_back(v);
                    touched.push_back(v);
                }
            }
        }
        depth = level[q.back()];
        far = q.back();
        for (int32_t v : q)
            if (level[v] == depth && (deg[v] < deg[far] || (deg[v] == deg[far] && v < far))) far = v;
        for (int32_t v : touched) level[v] = -1;
    };
    std::vector<int32_t> order;
    order.reserve(nc);
    for (int64_t seed = 0; seed < nc; seed++) {
        if (done[seed]) continue;
        int32_t s = (int32_t)seed, far, depth, d2;
        bfs_probe(s, far, depth);
        for (int it = 0; it < 4; it++) {
            int32_t far2;
            bfs_probe(far, far2, d2);
            if (d2 <= depth) break;
            s = far;
            far = far2;
            depth = d2;
        }
        // Cuthill-McKee from s
        size_t h = order.size();
        order.push_back(s);
        done[s] = 1;
        while (h < order.size()) {
            int32_t u = order[h++];
            for (int64_t j = ap[u]; j < ap[u + 1]; j++) {
                int32_t v = adj[j];
                if (!done[v]) {
                    done[v] = 1;
                    order.push_back(v);
                }
            }
        }
    }
    std::reverse(order.begin(), order.end());
    old_of_new = order;
}

// Hilbert index of a point with n = 2 or 3 coordinates of b bits each, by J. Skilling's transform
// ("Programming the Hilbert curve", AIP Conf. Proc. 707, 2004): undo the excess work of the axes
// (invert or exchange low bits from the top bit down), Gray-encode, then interleave the transposed
// coordinates most significant bit first.
uint64_t hilbert_index(uint32_t *X, int n, int b) {
    const uint32_t M = 1u << (b - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int i = 0; i < n; i++) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    for (int i = 1; i < n; i++) X[i] ^= X[i - 1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[n - 1] & Q) t ^= Q - 1;
    for (int i = 0; i < n; i++) X[i] ^= t;
    uint64_t key = 0;
    for (int bit = b - 1; bit >= 0; bit--)
        for (int i = 0; i < n; i++) key = (key << 1) | ((X[i] >> bit) & 1u);
    return key;
}

// Cells along the Hilbert curve through their centres (bounding box scaled to 2^b per axis, ties by
// file index): consecutive runs of cells are compact patches of the mesh, so most faces of a
// 256-cell block join two cells of that block.
void hilbert(const Prepared &P, std::vector<int32_t> &old_of_new) {
    const int64_t nc = P.nc;
    const int n = P.dim, b = n == 2 ? 31 : 21;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t c = 0; c < nc; c++)
        for (int d = 0; d < n; d++) {
            const double x = P.ccen[c * n + d];
            if (std::isfinite(x)) {
                lo[d] = std::min(lo[d], x);
                hi[d] = std::max(hi[d], x);
            }
        }
    double ext = 0;
    for (int d = 0; d < n; d++) ext = std::max(ext, hi[d] - lo[d]);
    const double top = (double)((1u << b) - 1);
    const double scale = ext > 0 ? top / ext : 0;
    std::vector<uint64_t> key(nc);
    for (int64_t c = 0; c < nc; c++) {
        uint32_t X[3] = {0, 0, 0};
        for (int d = 0; d < n; d++) {
            const double x = P.ccen[c * n + d];
            const double q = std::isfinite(x) ? std::min(std::max((x - lo[d]) * scale, 0.0), top) : 0.0;
            X[d] = (uint32_t)q;
        }
        key[c] = hilbert_index(X, n, b);
    }
    old_of_new.resize(nc);
    std::iota(old_of_new.begin(), old_of_new.end(), 0);
    std::sort(old_of_new.begin(), old_of_new.end(),
              [&](int32_t x, int32_t y) { return key[x] != key[y] ? key[x] < key[y] : x < y; });
}

int32_t kind_of(int32_t t) {
    switch (t) {
        case MST_BC_INTERIOR: return KIND_INTERIOR;
        case MST_BC_WALL: return KIND_WALL;
        case MST_BC_SYMMETRY: return KIND_SYMMETRY;
        case MST_BC_INLET: return KIND_INLET;
        case MST_BC_OUTLET: return KIND_OUTLET;
        case MST_BC_PERIODIC: return KIND_PERIODIC;
        default: return KIND_NONE;
    }
}

}  // namespace

int cell_centres(const mst_mesh &m, std::vector<double> &ccen) {
    int rc = validate(m);
    if (rc) return rc;
    Prepared P;
    std::vector<int64_t> cptr;
    std::vector<int32_t> cface;
    P.dim = m.dim;
    geometry(m, P, cptr, cface);
    ccen.swap(P.ccen);
    return MST_OK;
}

namespace {
std::atomic<int> g_cell_order{MST_ORDER_HILBERT};
}

int cell_order() { return g_cell_order.load(); }
void set_cell_order(int order) { g_cell_order.store(order); }

int prepare(const mst_mesh &m, Prepared &P) {
    int rc = validate(m);
    if (rc) return rc;
    const int dim = m.dim;
    const int64_t nf = m.n_faces, nc = m.n_cells;
    P = Prepared();
    P.dim = dim;
    P.nv = dim + 2;
    P.nc = nc;
    P.nf = nf;
    std::vector<int64_t> cptr;
    std::vector<int32_t> cface;
    geometry(m, P, cptr, cface);
    P.bad_cells = 0;
    for (int64_t c = 0; c < nc; c++)
        if (!(P.vol[c] > 0)) {
            if (P.bad_cells == 0) P.first_bad_cell = c;
            P.bad_cells++;
        }

    // ---- cells: Hilbert curve (default) or reverse Cuthill-McKee
    if (cell_order() == MST_ORDER_RCM)
        rcm(m, P.cell_old_of_new);
    else
        hilbert(P, P.cell_old_of_new);
    P.cell_new_of_old.assign(nc, -1);
    for (int64_t i = 0; i < nc; i++) P.cell_new_of_old[P.cell_old_of_new[i]] = (int32_t)i;

    // ---- faces: interior by (lowest new cell, highest new cell), then boundary kinds by new c0
    std::vector<int32_t> fo(nf);
    std::iota(fo.begin(), fo.end(), 0);
    auto key = [&](int32_t f, int64_t &k0, int64_t &k1, int64_t &k2) {
        int32_t a = P.cell_new_of_old[m.c0[f]];
        if (m.c1[f] >= 0) {
            int32_t b = P.cell_new_of_old[m.c1[f]];
            k0 = 0;
            k1 = std::min(a, b);
            k2 = std::max(a, b);
        } else {
            int32_t kd = kind_of(m.ftype[f]);
            if (kd == KIND_PERIODIC && m.partner[f] >= 0 && m.partner[f] < f) kd = KIND_MERGED;
            k0 = 1 + kd;
            k1 = a;
            k2 = 0;
        }
    };
    std::sort(fo.begin(), fo.end(), [&](int32_t x, int32_t y) {
        int64_t a0, a1, a2, b0, b1, b2;
        key(x, a0, a1, a2);
        key(y, b0, b1, b2);
        if (a0 != b0) return a0 < b0;
        if (a1 != b1) return a1 < b1;
        if (a2 != b2) return a2 < b2;
        return x < y;
    });
    P.face_old_of_new = fo;
    P.face_new_of_old.assign(nf, -1);
    for (int64_t i = 0; i < nf; i++) P.face_new_of_old[fo[i]] = (int32_t)i;

    // ---- SoA in new order
    P.fc0.resize(nf);
    P.fc1.resize(nf);
    P.fkind.resize(nf);
    P.ffile_branch2.resize(nf);
    P.fn.resize(nf * dim);
    P.fxf.resize(nf * dim);
    P.feta.resize(nf);
    P.n_interior = 0;
    P.pstart = -1;
    P.pcount = 0;
    for (int64_t i = 0; i < nf; i++) {
        int32_t f = fo[i];
        P.fc0[i] = P.cell_new_of_old[m.c0[f]];
        int32_t k = kind_of(m.ftype[f]);
        if (m.c1[f] >= 0) {
            P.fc1[i] = P.cell_new_of_old[m.c1[f]];
            k = KIND_INTERIOR;
            P.n_interior = i + 1;
        } else if (k == KIND_PERIODIC && m.partner[f] >= 0 && m.partner[f] > f) {
            // a periodic pair is one face: c0 on this side, c1 = the partner face's cell
            P.fc1[i] = P.cell_new_of_old[m.c0[m.partner[f]]];
            if (P.pstart < 0) P.pstart = i;
            P.pcount++;
        } else if (k == KIND_PERIODIC && m.partner[f] >= 0) {
            P.fc1[i] = -1;
            k = KIND_MERGED;
        } else {
            P.fc1[i] = -1;
            if (k == KIND_PERIODIC || k == KIND_INTERIOR) k = KIND_NONE;
        }
        P.fkind[i] = k;
        P.ffile_branch2[i] = P.dac[f] == 1 ? 1 : 0;
        for (int d = 0; d < dim; d++) {
            P.fn[i * dim + d] = (double)P.dac[f] * P.S[(int64_t)f * dim + d];
            P.fxf[i * dim + d] = P.fcen[(int64_t)f * dim + d];
        }
        P.feta[i] = P.eta[f];
    }
    if (P.pstart < 0) P.pstart = 0;
    if (P.pcount > 0 && P.pstart + P.pcount > nf) return fail(MST_E_MESH, "periodic faces not contiguous");
    P.fmerged_into.assign(nf, -1);
    for (int64_t i = 0; i < nf; i++)
        if (P.fkind[i] == KIND_MERGED) P.fmerged_into[i] = P.face_new_of_old[m.partner[fo[i]]];
    P.pxf.resize(P.pcount * dim);
    for (int64_t i = 0; i < P.pcount; i++) {
        int32_t f = fo[P.pstart + i];
        if (m.ftype[f] != MST_BC_PERIODIC) return fail(MST_E_MESH, "periodic faces not contiguous");
        int32_t g = m.partner[f];
        for (int d = 0; d < dim; d++) P.pxf[i * dim + d] = P.fcen[(int64_t)g * dim + d];
    }
    P.quirk_face = (m.n_int >= 1 && m.n_int <= nf && m.c1[m.n_int - 1] >= 0) ? P.face_new_of_old[m.n_int - 1] : -1;

    P.cvol.resize(nc);
    P.cxc.resize(nc * dim);
    std::vector<double> sh(nc * dim);
    for (int64_t i = 0; i < nc; i++) {
        int32_t c = P.cell_old_of_new[i];
        P.cvol[i] = P.vol[c];
        for (int d = 0; d < dim; d++) {
            P.cxc[i * dim + d] = P.ccen[(int64_t)c * dim + d];
            sh[i * dim + d] = P.cshalf[(int64_t)c * dim + d];
        }
    }
    P.cshalf.swap(sh);

    // ---- cell -> face slots (new ids, ascending); a periodic pair is listed on both of its cells
    // through its first face, the merged second face on none
    auto two_sided = [&](int64_t i) { return P.fkind[i] == KIND_INTERIOR || P.fkind[i] == KIND_PERIODIC; };
    std::vector<int32_t> cnt(nc, 0);
    for (int64_t i = 0; i < nf; i++) {
        if (P.fkind[i] == KIND_MERGED) continue;
        cnt[P.fc0[i]]++;
        if (two_sided(i)) cnt[P.fc1[i]]++;
    }
    P.maxf = 0;
    for (int64_t c = 0; c < nc; c++) P.maxf = std::max(P.maxf, cnt[c]);
    if (P.maxf > 8) return fail(MST_E_UNSUPPORTED, "cells with more than 8 faces are not supported");
    P.ell.assign((int64_t)P.maxf * nc, -1);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int64_t i = 0; i < nf; i++) {
        if (P.fkind[i] == KIND_MERGED) continue;
        int32_t a = P.fc0[i];
        P.ell[(int64_t)cnt[a]++ * nc + a] = (int32_t)(i << 1);
        if (two_sided(i)) {
            int32_t b = P.fc1[i];
            P.ell[(int64_t)cnt[b]++ * nc + b] = (int32_t)((i << 1) | 1);
        }
    }

    // ---- greedy face colouring: no two faces of one colour share a cell they accumulate into
    P.colour_new.assign(nf, -1);
    std::vector<uint32_t> used(nc, 0);
    P.ncolours = 0;
    for (int64_t i = 0; i < nf; i++) {
        uint32_t u = used[P.fc0[i]];
        if (two_sided(i)) u |= used[P.fc1[i]];
        int col = 0;
        while (col < 32 && (u >> col) & 1u) col++;
        if (col >= 32) return fail(MST_E_UNSUPPORTED, "face colouring needs more than 32 colours");
        P.colour_new[i] = col;
        used[P.fc0[i]] |= 1u << col;
        if (two_sided(i)) used[P.fc1[i]] |= 1u << col;
        P.ncolours = std::max(P.ncolours, col + 1);
    }
    return MST_OK;
}

int64_t incidence(const Prepared &P, int s, int64_t c, int32_t &code, double *n, double *dx) {
    const int dim = P.dim;
    code = -2;
    if (s >= P.maxf) return -1;
    const int32_t e = P.ell[(int64_t)s * P.nc + c];
    if (e < 0) return -1;
    const int64_t f = e >> 1;
    const bool is1 = e & 1;
    const int32_t nb = is1 ? P.fc0[f] : P.fc1[f];
    code = nb >= 0 ? (int32_t)(((int64_t)nb << 1) | (is1 ? 1 : 0)) : -1;
    const bool per = is1 && P.fkind[f] == KIND_PERIODIC;
    for (int d = 0; d < dim; d++) {
        n[d] = is1 ? -P.fn[f * dim + d] : P.fn[f * dim + d];
        const double xf = per ? P.pxf[(f - P.pstart) * dim + d] : P.fxf[f * dim + d];
        dx[d] = xf - P.cxc[c * dim + d];
    }
    return f;
}

void lr_masks(const Prepared &P, int compat, std::vector<int32_t> &mask) {
    const int dim = P.dim;
    mask.assign(P.nf, 0);
    for (int64_t i = 0; i < P.nf; i++) {
        int k = P.fkind[i];
        int32_t mk = 0;
        if (k == KIND_MERGED) {
            mk = 0;
        } else if (k == KIND_OUTLET) {
            mk = (1 << dim) - 1;  // outlet: set(Q_c0 side first) in both directions (RhoSolver.cpp:219-224,358-361)
        } else if (k != KIND_NONE) {
            for (int d = 0; d < dim; d++) {
                // flag_d = (directAndCells * S_d >= 0): MshBlock.cpp:286; inverted on branch-2 faces under compat
                int phys = P.fn[i * dim + d] >= 0;
                if ((compat & MST_COMPAT_BRANCH2_LR) && dim == 2 && P.ffile_branch2[i]) phys = !phys;
                if (phys) mk |= 1 << d;
            }
        }
        mask[i] = mk;
    }
}

// LU-SGS colouring of the cell graph (cells sharing a face get different colours).  Fewer colours
// mean fewer sweep launches per iteration (2 ncol - 1), each of which costs a fixed few microseconds:
//  1. two colours when the graph is bipartite (breadth-first search from each uncoloured cell in
//     device order; the split-quad meshes are);
//  2. otherwise greedy in device order (each cell takes the smallest colour none of its
//     lower-numbered neighbours has; at most maxf + 1 colours), then Kempe-chain swaps that empty
//     the top colour where they can: a top-colour cell whose neighbours use colours a and b swaps a
//     and b along the {a, b}-chain through its a-neighbour; if that chain misses its b-neighbours,
//     colour a becomes free for it.  Chains are searched up to a fixed size.
// Deterministic.  Returns the number of colours (every colour 0..n-1 is used).
static void cell_neighbours(const Prepared &P, int64_t c, int32_t (&nb)[8], int &n) {
    n = 0;
    for (int s = 0; s < P.maxf; s++) {
        const int32_t e = P.ell[(int64_t)s * P.nc + c];
        if (e < 0) break;
        const int32_t f = e >> 1;
        const int32_t o = (e & 1) ? P.fc0[f] : P.fc1[f];
        if (o >= 0 && o != c) nb[n++] = o;
    }
}

static bool two_colours(const Prepared &P, std::vector<int32_t> &col) {
    const int64_t nc = P.nc;
    col.assign(nc, -1);
    std::vector<int32_t> queue;
    queue.reserve(nc);
    for (int64_t s0 = 0; s0 < nc; s0++) {
        if (col[s0] >= 0) continue;
        col[s0] = 0;
        queue.clear();
        queue.push_back((int32_t)s0);
        for (size_t h = 0; h < queue.size(); h++) {
            const int32_t c = queue[h];
            int32_t nb[8];
            int n;
            cell_neighbours(P, c, nb, n);
            for (int j = 0; j < n; j++) {
                if (col[nb[j]] < 0) {
                    col[nb[j]] = 1 - col[c];
                    queue.push_back(nb[j]);
                } else if (col[nb[j]] == col[c]) {
                    return false;
                }
            }
        }
    }
    return true;
}

static void greedy_kempe(const Prepared &P, std::vector<int32_t> &col) {
    const int64_t nc = P.nc;
    col.assign(nc, -1);
    int ncol = 0;
    for (int64_t c = 0; c < nc; c++) {
        int32_t nb[8];
        int n;
        cell_neighbours(P, c, nb, n);
        uint32_t used = 0;
        for (int j = 0; j < n; j++)
            if (col[nb[j]] >= 0) used |= 1u << col[nb[j]];
        int k = 0;
        while ((used >> k) & 1u) k++;
        col[c] = k;
        ncol = std::max(ncol, k + 1);
    }
    constexpr size_t CHAIN_MAX = 1 << 14;
    std::vector<int32_t> stamp(nc, -1), chain;
    int32_t pass = 0;
    while (ncol > 2) {
        const int top = ncol - 1;
        bool emptied = true;
        for (int64_t v = 0; v < nc; v++) {
            if (col[v] != top) continue;
            int32_t nb[8];
            int n;
            cell_neighbours(P, v, nb, n);
            bool fixed = false;
            for (int a = 0; a < top && !fixed; a++) {
                uint32_t used = 0;
                for (int j = 0; j < n; j++) used |= 1u << col[nb[j]];
                if (!((used >> a) & 1u)) {  // already free
                    col[v] = a;
                    fixed = true;
                    break;
                }
                for (int b = 0; b < top && !fixed; b++) {
                    if (b == a) continue;
                    int ua = -1;
                    for (int j = 0; j < n; j++)
                        if (col[nb[j]] == a) ua = nb[j];
                    // chain of colours {a, b} through one a-neighbour (every a-neighbour must be in it)
                    ++pass;
                    chain.clear();
                    chain.push_back(ua);
                    stamp[ua] = pass;
                    bool ok = true;
                    for (size_t h = 0; h < chain.size() && ok; h++) {
                        int32_t nb2[8];
                        int n2;
                        cell_neighbours(P, chain[h], nb2, n2);
                        for (int j = 0; j < n2; j++) {
                            const int32_t y = nb2[j];
                            if (stamp[y] != pass && (col[y] == a || col[y] == b)) {
                                stamp[y] = pass;
                                chain.push_back(y);
                            }
                        }
                        if (chain.size() > CHAIN_MAX) ok = false;
                    }
                    if (!ok) continue;
                    bool blocked = false;
                    for (int j = 0; j < n; j++)
                        if (col[nb[j]] == a && stamp[nb[j]] != pass) blocked = true;  // an a-neighbour outside
                        else if (col[nb[j]] == b && stamp[nb[j]] == pass) blocked = true;  // a b-neighbour inside
                    if (blocked) continue;
                    for (int32_t x : chain) col[x] = col[x] == a ? b : a;
                    col[v] = a;
                    fixed = true;
                }
            }
            if (!fixed) emptied = false;
        }
        if (!emptied) break;
        ncol--;
    }
}

int lusgs_colours(const Prepared &P, std::vector<int32_t> &col) {
    if (!two_colours(P, col)) greedy_kempe(P, col);
    return P.nc > 0 ? *std::max_element(col.begin(), col.end()) + 1 : 0;
}

}  // namespace mst
