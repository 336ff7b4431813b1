// capi_mesh.cpp -- C ABI of the host mesh and preprocessing (no device code; see include/mstcfd.h).
#include <cstring>
#include <new>

#include "mesh.hpp"
#include "prep.hpp"

namespace mst {
const std::string &last_error();
}

#define MST_GUARD(expr)                                                    \
    try {                                                                  \
        expr;                                                              \
    } catch (const std::bad_alloc &) {                                     \
        return mst::fail(MST_E_ARG, "out of host memory");                 \
    } catch (const std::exception &e) {                                    \
        return mst::fail(MST_E_ARG, std::string("host exception: ") + e.what()); \
    }

extern "C" {

size_t mst_last_error(char *buf, size_t n) {
    const std::string &e = mst::last_error();
    if (buf && n) {
        size_t k = e.size() < n - 1 ? e.size() : n - 1;
        memcpy(buf, e.data(), k);
        buf[k] = 0;
    }
    return e.size();
}

int mst_mesh_read_msh(const char *path, mst_mesh **out) {
    if (!path || !out) return mst::fail(MST_E_ARG, "null argument");
    *out = nullptr;
    MST_GUARD({
        mst_mesh *m = new mst_mesh();
        int rc = mst::read_msh(path, *m);
        if (rc) {
            delete m;
            return rc;
        }
        *out = m;
    });
    return MST_OK;
}

int mst_mesh_write_msh(const mst_mesh *m, const char *path) {
    if (!m || !path) return mst::fail(MST_E_ARG, "null argument");
    MST_GUARD(return mst::write_msh(*m, path));
}

int mst_mesh_write_cache(const mst_mesh *m, const char *path) {
    if (!m || !path) return mst::fail(MST_E_ARG, "null argument");
    MST_GUARD(return mst::write_mesh_cache(*m, path));
}

int mst_mesh_read_cache(const char *path, mst_mesh **out) {
    if (!path || !out) return mst::fail(MST_E_ARG, "null argument");
    *out = nullptr;
    MST_GUARD({
        mst_mesh *m = new mst_mesh();
        int rc = mst::read_mesh_cache(path, *m);
        if (rc) {
            delete m;
            return rc;
        }
        *out = m;
    });
    return MST_OK;
}

int mst_write_tecplot(const mst_mesh *m, const double *Q, int64_t n_cells, double gamma, double cv, int wall_noslip,
                      const char *title, const char *path) {
    if (!m || !Q || !path) return mst::fail(MST_E_ARG, "null argument");
    if (n_cells != m->n_cells) return mst::fail(MST_E_ARG, "size does not match the mesh");
    if (!(gamma > 1) || !(cv > 0)) return mst::fail(MST_E_ARG, "gamma must exceed 1 and cv be positive");
    MST_GUARD(return mst::write_tecplot(*m, Q, gamma, cv, wall_noslip, title ? title : "mst-cfd", path));
}

int mst_checkpoint_write(const char *path, const mst_mesh *m, const double *Q, int64_t n_cells, int64_t step,
                         double time) {
    if (!m || !Q || !path) return mst::fail(MST_E_ARG, "null argument");
    if (n_cells != m->n_cells) return mst::fail(MST_E_ARG, "size does not match the mesh");
    MST_GUARD(return mst::write_checkpoint(path, *m, Q, step, time));
}

int mst_checkpoint_read(const char *path, const mst_mesh *m, double *Q, int64_t n_cells, int64_t *step, double *time) {
    if (!m || !Q || !path) return mst::fail(MST_E_ARG, "null argument");
    if (n_cells != m->n_cells) return mst::fail(MST_E_ARG, "size does not match the mesh");
    MST_GUARD(return mst::read_checkpoint(path, *m, Q, step, time));
}

int mst_mesh_from_arrays(int dim, int64_t n_nodes, const double *xyz, int64_t n_cells, int64_t n_faces, int fnn,
                         const int32_t *fnodes, const int32_t *c0, const int32_t *c1, const int32_t *ftype,
                         const int32_t *partner, mst_mesh **out) {
    if (!out || !xyz || !fnodes || !c0 || !c1 || !ftype) return mst::fail(MST_E_ARG, "null argument");
    if (dim != 2 && dim != 3) return mst::fail(MST_E_ARG, "dim must be 2 or 3");
    if (n_nodes <= 0 || n_cells <= 0 || n_faces <= 0 || fnn < 2 || fnn > 4)
        return mst::fail(MST_E_ARG, "bad mesh sizes");
    *out = nullptr;
    MST_GUARD({
        mst_mesh *m = new mst_mesh();
        m->dim = dim;
        m->fnn = fnn;
        m->n_nodes = n_nodes;
        m->n_cells = n_cells;
        m->n_faces = n_faces;
        m->xyz.assign(xyz, xyz + n_nodes * dim);
        m->fnodes.assign(fnodes, fnodes + n_faces * fnn);
        m->c0.assign(c0, c0 + n_faces);
        m->c1.assign(c1, c1 + n_faces);
        m->ftype.assign(ftype, ftype + n_faces);
        if (partner)
            m->partner.assign(partner, partner + n_faces);
        else
            m->partner.assign(n_faces, -1);
        mst::rebuild_zones(*m);
        int rc = mst::validate(*m);
        if (rc) {
            delete m;
            return rc;
        }
        *out = m;
    });
    return MST_OK;
}

int mst_mesh_counts(const mst_mesh *m, int64_t out[7]) {
    if (!m || !out) return mst::fail(MST_E_ARG, "null argument");
    out[0] = m->dim;
    out[1] = m->fnn;
    out[2] = m->n_nodes;
    out[3] = m->n_cells;
    out[4] = m->n_faces;
    out[5] = m->n_int;
    out[6] = (int64_t)m->zones.size();
    return MST_OK;
}

int mst_mesh_arrays(const mst_mesh *m, double *xyz, int32_t *fnodes, int32_t *c0, int32_t *c1, int32_t *ftype,
                    int32_t *partner) {
    if (!m) return mst::fail(MST_E_ARG, "null argument");
    if (xyz) memcpy(xyz, m->xyz.data(), sizeof(double) * m->xyz.size());
    if (fnodes) memcpy(fnodes, m->fnodes.data(), sizeof(int32_t) * m->fnodes.size());
    if (c0) memcpy(c0, m->c0.data(), sizeof(int32_t) * m->c0.size());
    if (c1) memcpy(c1, m->c1.data(), sizeof(int32_t) * m->c1.size());
    if (ftype) memcpy(ftype, m->ftype.data(), sizeof(int32_t) * m->ftype.size());
    if (partner) memcpy(partner, m->partner.data(), sizeof(int32_t) * m->partner.size());
    return MST_OK;
}

int mst_mesh_zones(const mst_mesh *m, int32_t *out) {
    if (!m || !out) return mst::fail(MST_E_ARG, "null argument");
    for (size_t i = 0; i < m->zones.size(); i++) {
        out[4 * i + 0] = m->zones[i].id;
        out[4 * i + 1] = m->zones[i].first;
        out[4 * i + 2] = m->zones[i].last;
        out[4 * i + 3] = m->zones[i].type;
    }
    return MST_OK;
}

void mst_mesh_destroy(mst_mesh *m) { delete m; }

int mst_mesh_gen_tri(int nx, int ny, double x0, double x1, double y0, double y1, double jitter, uint64_t seed,
                     const int32_t bc[4], mst_mesh **out) {
    if (!out || !bc) return mst::fail(MST_E_ARG, "null argument");
    *out = nullptr;
    MST_GUARD({
        mst_mesh *m = new mst_mesh();
        int rc = mst::gen_tri(nx, ny, x0, x1, y0, y1, jitter, seed, bc, *m);
        if (rc) {
            delete m;
            return rc;
        }
        *out = m;
    });
    return MST_OK;
}

int mst_mesh_gen_ramp(int nx, int ny, double x0, double x1, double height, double theta_deg, double jitter,
                      uint64_t seed, const int32_t bc[4], mst_mesh **out) {
    if (!out || !bc) return mst::fail(MST_E_ARG, "null argument");
    *out = nullptr;
    MST_GUARD({
        mst_mesh *m = new mst_mesh();
        int rc = mst::gen_ramp(nx, ny, x0, x1, height, theta_deg, jitter, seed, bc, *m);
        if (rc) {
            delete m;
            return rc;
        }
        *out = m;
    });
    return MST_OK;
}

int mst_mesh_cell_centres(const mst_mesh *m, double *ccen, int64_t n_cells) {
    if (!m || !ccen) return mst::fail(MST_E_ARG, "null argument");
    if (n_cells != m->n_cells) return mst::fail(MST_E_ARG, "size does not match the mesh");
    MST_GUARD({
        std::vector<double> c;
        int rc = mst::cell_centres(*m, c);
        if (rc) return rc;
        std::copy(c.begin(), c.end(), ccen);
    });
    return MST_OK;
}

int mst_mesh_gen_tet(int nx, int ny, int nz, const double lo[3], const double hi[3], double jitter, uint64_t seed,
                     const int32_t bc[6], mst_mesh **out) {
    if (!out || !bc || !lo || !hi) return mst::fail(MST_E_ARG, "null argument");
    *out = nullptr;
    MST_GUARD({
        mst_mesh *m = new mst_mesh();
        int rc = mst::gen_tet(nx, ny, nz, lo, hi, jitter, seed, bc, *m);
        if (rc) {
            delete m;
            return rc;
        }
        *out = m;
    });
    return MST_OK;
}

int mst_set_cell_order(int order) {
    if (order != MST_ORDER_HILBERT && order != MST_ORDER_RCM) return mst::fail(MST_E_ARG, "bad cell order");
    mst::set_cell_order(order);
    return MST_OK;
}

int mst_get_cell_order(int *order) {
    if (!order) return mst::fail(MST_E_ARG, "null argument");
    *order = mst::cell_order();
    return MST_OK;
}

int mst_prep_create(const mst_mesh *m, mst_prep **out) {
    if (!m || !out) return mst::fail(MST_E_ARG, "null argument");
    *out = nullptr;
    MST_GUARD({
        mst_prep *p = new mst_prep();
        int rc = mst::prepare(*m, p->P);
        if (rc) {
            delete p;
            return rc;
        }
        *out = p;
    });
    return MST_OK;
}

void mst_prep_destroy(mst_prep *p) { delete p; }

int mst_prep_counts(const mst_prep *p, int64_t out[8]) {
    if (!p || !out) return mst::fail(MST_E_ARG, "null argument");
    const mst::Prepared &P = p->P;
    out[0] = P.dim;
    out[1] = P.nv;
    out[2] = P.nc;
    out[3] = P.nf;
    out[4] = P.maxf;
    out[5] = P.ncolours;
    out[6] = P.n_interior;
    out[7] = P.pcount;
    return MST_OK;
}

int mst_prep_permutation(const mst_prep *p, int32_t *cell_new_of_old, int32_t *face_new_of_old,
                         int32_t *face_colour) {
    if (!p) return mst::fail(MST_E_ARG, "null argument");
    const mst::Prepared &P = p->P;
    if (cell_new_of_old) memcpy(cell_new_of_old, P.cell_new_of_old.data(), sizeof(int32_t) * P.nc);
    if (face_new_of_old) memcpy(face_new_of_old, P.face_new_of_old.data(), sizeof(int32_t) * P.nf);
    if (face_colour)
        for (int64_t f = 0; f < P.nf; f++) face_colour[f] = P.colour_new[P.face_new_of_old[f]];
    return MST_OK;
}

int mst_prep_geometry(const mst_prep *p, double *fcen, double *S, double *eta, int32_t *dac, double *ccen,
                      double *vol) {
    if (!p) return mst::fail(MST_E_ARG, "null argument");
    const mst::Prepared &P = p->P;
    if (fcen) memcpy(fcen, P.fcen.data(), sizeof(double) * P.fcen.size());
    if (S) memcpy(S, P.S.data(), sizeof(double) * P.S.size());
    if (eta) memcpy(eta, P.eta.data(), sizeof(double) * P.eta.size());
    if (dac) memcpy(dac, P.dac.data(), sizeof(int32_t) * P.dac.size());
    if (ccen) memcpy(ccen, P.ccen.data(), sizeof(double) * P.ccen.size());
    if (vol) memcpy(vol, P.vol.data(), sizeof(double) * P.vol.size());
    return MST_OK;
}

int mst_prep_tables(const mst_prep *p, int32_t *fc0, int32_t *fc1, int32_t *kind, int32_t *cell_faces) {
    if (!p) return mst::fail(MST_E_ARG, "null argument");
    const mst::Prepared &P = p->P;
    if (fc0) memcpy(fc0, P.fc0.data(), sizeof(int32_t) * P.nf);
    if (fc1) memcpy(fc1, P.fc1.data(), sizeof(int32_t) * P.nf);
    if (kind) memcpy(kind, P.fkind.data(), sizeof(int32_t) * P.nf);
    if (cell_faces) memcpy(cell_faces, P.ell.data(), sizeof(int32_t) * P.ell.size());
    return MST_OK;
}

int mst_prep_cell_colours(const mst_prep *p, int32_t *colour, int *n_colours) {
    if (!p || !colour || !n_colours) return mst::fail(MST_E_ARG, "null argument");
    const mst::Prepared &P = p->P;
    std::vector<int32_t> col;
    *n_colours = mst::lusgs_colours(P, col);
    for (int64_t c = 0; c < P.nc; c++) colour[c] = col[P.cell_new_of_old[c]];
    return MST_OK;
}

}  // extern "C"
