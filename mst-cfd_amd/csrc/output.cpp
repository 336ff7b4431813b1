// output.cpp -- on-disk formats around the residual path: the Tecplot ASCII result file and a binary
// checkpoint for restarts.  Host code, file order throughout (the caller's state arrays).
//
// Tecplot writer: Work::writedataRhoBasedMshNodePlt (work/Work.cpp:204-304).  Cell states are first
// carried to the faces (interior: eta0-weighted mean of the two cells; boundary: the cell's value,
// walls optionally with zero velocity as FLAG_WALLSMOOTH = 0 does, :246-285), then to the nodes (mean
// of the adjacent faces' values weighted by 1 / face area, :286-299), and written as
// X Y [Z] rho u v [w] T p Ma per node plus the cells' node lists (:300-306).  Differences, each a fix:
//  - the zone type follows the mesh (all triangles, triangles and quadrilaterals -- a triangle
//    repeats its last node --, or tetrahedra) instead of the FELNUM macro (:231-235);
//  - a node's face weights are those of its own faces (:292-293 indexes the face array by the node
//    number);
//  - symmetry and periodic faces get values (the reference's switch leaves them unset, :246-285);
//  - the TITLE line is well formed (:227 opens it with a stray quote) and 3D files list w
//    (:238 lists three velocities' worth of columns under two names);
//  - numbers are written in scientific notation with 17 significant digits (round-trip exact)
//    instead of fixed-point with 15 decimals (which loses small values).
//
// Checkpoint: replaces the text .cellPUVT dump (Work::writedataCellContinue / readPbasedCellPlt,
// work/Work.cpp:432-514) with a binary file that restores the state bit for bit and refuses a file
// written for another mesh or damaged in between:
//   header  "MSTCKPT1", version, dim, nv, n_cells, n_faces, n_nodes, mesh fingerprint, step, time,
//           state checksum (FNV-1a 64 of the state's bytes)
//   data    n_cells x nv doubles, the state in file order
// in the host's byte order (little-endian on x86-64).
// The fingerprint is FNV-1a 64 over the mesh's counts, node coordinates and face tables.
//
// Mesh cache: the reader's result (mst_mesh, file order) in binary, so a large mesh (C3's Fluent text
// runs to gigabytes) is parsed once:
//   "MSTMESH1", version, dim, fnn, counts (nodes, cells, faces, n_int, zones), the arrays
//   (xyz, fnodes, c0, c1, ftype, partner), each zone (id, first, last, type, name), and an FNV-1a 64
//   checksum of everything before it.  Reading validates the mesh as the reader does.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mesh.hpp"
#include "prep.hpp"

namespace mst {
namespace {

struct Fnv {
    uint64_t h = 1469598103934665603ull;
    void add(const void *p, size_t n) {
        const unsigned char *b = (const unsigned char *)p;
        for (size_t i = 0; i < n; i++) {
            h ^= b[i];
            h *= 1099511628211ull;
        }
    }
    template <typename T>
    void add_vec(const std::vector<T> &v) {
        add(v.data(), v.size() * sizeof(T));
    }
};

uint64_t mesh_fingerprint(const mst_mesh &m) {
    Fnv f;
    const int64_t counts[5] = {m.dim, m.fnn, m.n_nodes, m.n_cells, m.n_faces};
    f.add(counts, sizeof counts);
    f.add_vec(m.xyz);
    f.add_vec(m.fnodes);
    f.add_vec(m.c0);
    f.add_vec(m.c1);
    f.add_vec(m.ftype);
    f.add_vec(m.partner);
    return f.h;
}

struct CkptHeader {
    char magic[8];
    uint32_t version, dim, nv, reserved;
    int64_t n_cells, n_faces, n_nodes;
    uint64_t mesh_hash;
    int64_t step;
    double time;
    uint64_t data_hash;
};
const char CKPT_MAGIC[8] = {'M', 'S', 'T', 'C', 'K', 'P', 'T', '1'};

// the cell's nodes: a triangle or tetrahedron in any order, a quadrilateral in cyclic order; the node
// count, or -1 for another cell shape
int cell_nodes(const mst_mesh &m, const std::vector<std::vector<int32_t>> &cfaces, int64_t c, int32_t (&out)[4]) {
    int n = 0;
    for (int32_t f : cfaces[c])
        for (int k = 0; k < m.fnn; k++) {
            const int32_t v = m.fnodes[(int64_t)f * m.fnn + k];
            bool seen = false;
            for (int j = 0; j < n; j++) seen |= out[j] == v;
            if (!seen) {
                if (n == 4) return -1;
                out[n++] = v;
            }
        }
    if (m.dim == 3) return n == 4 && cfaces[c].size() == 4 ? 4 : -1;
    if (n == 3) return 3;
    if (n != 4 || cfaces[c].size() != 4) return -1;
    // quadrilateral: walk the edges from the first face's two nodes
    int32_t ring[4];
    ring[0] = m.fnodes[(int64_t)cfaces[c][0] * 2];
    ring[1] = m.fnodes[(int64_t)cfaces[c][0] * 2 + 1];
    std::vector<char> used(cfaces[c].size(), 0);
    used[0] = 1;
    for (int pos = 2; pos < 4; pos++) {
        bool found = false;
        for (size_t j = 0; j < cfaces[c].size() && !found; j++) {
            if (used[j]) continue;
            const int32_t a = m.fnodes[(int64_t)cfaces[c][j] * 2], b = m.fnodes[(int64_t)cfaces[c][j] * 2 + 1];
            if (a == ring[pos - 1] || b == ring[pos - 1]) {
                ring[pos] = a == ring[pos - 1] ? b : a;
                used[j] = 1;
                found = true;
            }
        }
        if (!found) return -1;
    }
    for (int k = 0; k < 4; k++) out[k] = ring[k];
    return 4;
}

}  // namespace

int write_tecplot(const mst_mesh &m, const double *Q, double gamma, double cv, int wall_noslip, const std::string &title,
                  const std::string &path) {
    const int dim = m.dim, nv = dim + 2;
    Prepared P;
    int rc = prepare(m, P);  // file-order geometry: eta0 and area vectors (the reference's Face data)
    if (rc) return rc;
    std::vector<std::vector<int32_t>> cfaces(m.n_cells);
    for (int64_t f = 0; f < m.n_faces; f++) {
        cfaces[m.c0[f]].push_back((int32_t)f);
        if (m.c1[f] >= 0) cfaces[m.c1[f]].push_back((int32_t)f);
    }
    std::vector<int32_t> conn((size_t)m.n_cells * 4);
    bool all_tri = true;
    for (int64_t c = 0; c < m.n_cells; c++) {
        int32_t nodes[4];
        const int n = cell_nodes(m, cfaces, c, nodes);
        if (n < 0)
            return fail(MST_E_UNSUPPORTED, "Tecplot output covers triangles, quadrilaterals and tetrahedra only (cell " +
                                               std::to_string(c) + ")");
        for (int k = 0; k < 4; k++) conn[(size_t)c * 4 + k] = nodes[k < n ? k : n - 1];
        all_tri &= n == 3;
    }
    // face values (Work.cpp:246-285)
    std::vector<double> fq((size_t)m.n_faces * nv), w(m.n_faces);
    for (int64_t f = 0; f < m.n_faces; f++) {
        const double *q0 = Q + (int64_t)m.c0[f] * nv;
        double *o = &fq[(size_t)f * nv];
        const double e = P.eta[f];
        int64_t other = -1;
        if (m.c1[f] >= 0)
            other = m.c1[f];
        else if (m.ftype[f] == MST_BC_PERIODIC && m.partner[f] >= 0)
            other = m.c0[m.partner[f]];
        if (other >= 0) {
            const double *q1 = Q + other * nv;
            for (int k = 0; k < nv; k++) o[k] = e * q0[k] + (1 - e) * q1[k];
        } else {
            for (int k = 0; k < nv; k++) o[k] = q0[k];
            if (m.ftype[f] == MST_BC_WALL && wall_noslip)
                for (int d = 0; d < dim; d++) o[1 + d] = 0;
        }
        double a2 = 0;
        for (int d = 0; d < dim; d++) a2 += P.S[(size_t)f * dim + d] * P.S[(size_t)f * dim + d];
        w[f] = 1.0 / std::sqrt(a2);
    }
    // node values: 1/area-weighted mean over the node's faces (Work.cpp:286-299)
    std::vector<double> nq((size_t)m.n_nodes * nv, 0.0), nw(m.n_nodes, 0.0);
    for (int64_t f = 0; f < m.n_faces; f++)
        for (int k = 0; k < m.fnn; k++) {
            const int32_t v = m.fnodes[(int64_t)f * m.fnn + k];
            for (int j = 0; j < nv; j++) nq[(size_t)v * nv + j] += w[f] * fq[(size_t)f * nv + j];
            nw[v] += w[f];
        }
    FILE *fp = fopen(path.c_str(), "w");
    if (!fp) return fail(MST_E_IO, "cannot open " + path + " for writing");
    static const char *vel[3] = {"u", "v", "w"}, *crd[3] = {"X", "Y", "Z"};
    fprintf(fp, "TITLE = \"%s\"\nVARIABLES =", title.c_str());
    for (int d = 0; d < dim; d++) fprintf(fp, " \"%s\",", crd[d]);
    fprintf(fp, " \"rho\",");
    for (int d = 0; d < dim; d++) fprintf(fp, " \"%s\",", vel[d]);
    fprintf(fp, " \"T\", \"p\", \"Ma\"\n");
    const char *zt = dim == 3 ? "FETETRAHEDRON" : (all_tri ? "FETRIANGLE" : "FEQUADRILATERAL");
    fprintf(fp, "ZONE T=\"%s\", DATAPACKING=POINT, NODES=%lld, ELEMENTS=%lld, ZONETYPE=%s\n", title.c_str(),
            (long long)m.n_nodes, (long long)m.n_cells, zt);
    for (int64_t v = 0; v < m.n_nodes; v++) {
        double q[5];
        for (int j = 0; j < nv; j++) q[j] = nw[v] > 0 ? nq[(size_t)v * nv + j] / nw[v] : 0.0;
        double ke = 0;
        for (int d = 0; d < dim; d++) ke += q[1 + d] * q[1 + d];
        const double e = q[nv - 1] - 0.5 * ke / q[0];  // internal energy per volume (work/FUNCTION.cpp:8-20)
        const double T = e / q[0] / cv, p = e * (gamma - 1), ma = std::sqrt(ke / (gamma * p * q[0]));
        for (int d = 0; d < dim; d++) fprintf(fp, "%.17g ", m.xyz[(size_t)v * dim + d]);
        fprintf(fp, "%.17g", q[0]);
        for (int d = 0; d < dim; d++) fprintf(fp, " %.17g", q[1 + d] / q[0]);
        fprintf(fp, " %.17g %.17g %.17g\n", T, p, ma);
    }
    const int per = dim == 3 ? 4 : (all_tri ? 3 : 4);
    for (int64_t c = 0; c < m.n_cells; c++) {
        for (int k = 0; k < per; k++) fprintf(fp, k ? " %d" : "%d", conn[(size_t)c * 4 + k] + 1);
        fputc('\n', fp);
    }
    const bool ok = !ferror(fp);
    if (fclose(fp) != 0 || !ok) return fail(MST_E_IO, "write failed: " + path);
    return MST_OK;
}

int write_checkpoint(const std::string &path, const mst_mesh &m, const double *Q, int64_t step, double time) {
    CkptHeader h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, CKPT_MAGIC, 8);
    h.version = 1;
    h.dim = (uint32_t)m.dim;
    h.nv = (uint32_t)(m.dim + 2);
    h.n_cells = m.n_cells;
    h.n_faces = m.n_faces;
    h.n_nodes = m.n_nodes;
    h.mesh_hash = mesh_fingerprint(m);
    h.step = step;
    h.time = time;
    const size_t nbytes = sizeof(double) * (size_t)m.n_cells * h.nv;
    Fnv d;
    d.add(Q, nbytes);
    h.data_hash = d.h;
    const std::string tmp = path + ".tmp";
    FILE *fp = fopen(tmp.c_str(), "wb");
    if (!fp) return fail(MST_E_IO, "cannot open " + tmp + " for writing");
    bool ok = fwrite(&h, sizeof h, 1, fp) == 1 && fwrite(Q, 1, nbytes, fp) == nbytes;
    ok = (fclose(fp) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        return fail(MST_E_IO, "write failed: " + path);
    }
    return MST_OK;
}

int read_checkpoint(const std::string &path, const mst_mesh &m, double *Q, int64_t *step, double *time) {
    FILE *fp = fopen(path.c_str(), "rb");
    if (!fp) return fail(MST_E_IO, "cannot open " + path);
    CkptHeader h;
    const bool got = fread(&h, sizeof h, 1, fp) == 1;
    if (!got || memcmp(h.magic, CKPT_MAGIC, 8) != 0 || h.version != 1 || h.nv != (uint32_t)(m.dim + 2)) {
        fclose(fp);
        return fail(MST_E_IO, path + ": not a checkpoint of this format");
    }
    if ((int)h.dim != m.dim || h.n_cells != m.n_cells || h.n_faces != m.n_faces || h.n_nodes != m.n_nodes ||
        h.mesh_hash != mesh_fingerprint(m)) {
        fclose(fp);
        return fail(MST_E_ARG, path + ": written for another mesh");
    }
    const size_t nbytes = sizeof(double) * (size_t)m.n_cells * h.nv;
    std::vector<double> buf((size_t)m.n_cells * h.nv);
    const bool full = fread(buf.data(), 1, nbytes, fp) == nbytes;
    const bool at_end = fgetc(fp) == EOF;
    fclose(fp);
    if (!full || !at_end) return fail(MST_E_IO, path + ": truncated or trailing data");
    Fnv d;
    d.add(buf.data(), nbytes);
    if (d.h != h.data_hash) return fail(MST_E_IO, path + ": state checksum mismatch (damaged file)");
    memcpy(Q, buf.data(), nbytes);
    if (step) *step = h.step;
    if (time) *time = h.time;
    return MST_OK;
}

namespace {
const char MESH_MAGIC[8] = {'M', 'S', 'T', 'M', 'E', 'S', 'H', '1'};

struct Writer {
    std::vector<unsigned char> buf;
    void put(const void *p, size_t n) {
        const unsigned char *b = (const unsigned char *)p;
        buf.insert(buf.end(), b, b + n);
    }
    template <typename T>
    void val(T v) { put(&v, sizeof v); }
    template <typename T>
    void vec(const std::vector<T> &v) { put(v.data(), v.size() * sizeof(T)); }
};

struct Reader {
    const unsigned char *p, *end;
    bool ok = true;
    void get(void *dst, size_t n) {
        if ((size_t)(end - p) < n) {
            ok = false;
            return;
        }
        memcpy(dst, p, n);
        p += n;
    }
    template <typename T>
    T val() {
        T v{};
        get(&v, sizeof v);
        return v;
    }
    template <typename T>
    void vec(std::vector<T> &v, int64_t n) {
        if (n < 0 || (uint64_t)n > (uint64_t)(end - p) / sizeof(T)) {
            ok = false;
            return;
        }
        v.resize((size_t)n);
        get(v.data(), (size_t)n * sizeof(T));
    }
};
}  // namespace

int write_mesh_cache(const mst_mesh &m, const std::string &path) {
    Writer w;
    w.put(MESH_MAGIC, 8);
    w.val<uint32_t>(1);
    w.val<int32_t>(m.dim);
    w.val<int32_t>(m.fnn);
    w.val<int64_t>(m.n_nodes);
    w.val<int64_t>(m.n_cells);
    w.val<int64_t>(m.n_faces);
    w.val<int64_t>(m.n_int);
    w.val<int64_t>((int64_t)m.zones.size());
    w.vec(m.xyz);
    w.vec(m.fnodes);
    w.vec(m.c0);
    w.vec(m.c1);
    w.vec(m.ftype);
    w.vec(m.partner);
    for (const mst_zone &z : m.zones) {
        w.val<int32_t>(z.id);
        w.val<int32_t>(z.first);
        w.val<int32_t>(z.last);
        w.val<int32_t>(z.type);
        w.val<int32_t>((int32_t)z.name.size());
        w.put(z.name.data(), z.name.size());
    }
    Fnv h;
    h.add(w.buf.data(), w.buf.size());
    w.val<uint64_t>(h.h);
    const std::string tmp = path + ".tmp";
    FILE *fp = fopen(tmp.c_str(), "wb");
    if (!fp) return fail(MST_E_IO, "cannot open " + tmp + " for writing");
    bool ok = fwrite(w.buf.data(), 1, w.buf.size(), fp) == w.buf.size();
    ok = (fclose(fp) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        return fail(MST_E_IO, "write failed: " + path);
    }
    return MST_OK;
}

int read_mesh_cache(const std::string &path, mst_mesh &m) {
    FILE *fp = fopen(path.c_str(), "rb");
    if (!fp) return fail(MST_E_IO, "cannot open " + path);
    std::vector<unsigned char> buf;
    unsigned char chunk[1 << 16];
    size_t got;
    while ((got = fread(chunk, 1, sizeof chunk, fp)) > 0) buf.insert(buf.end(), chunk, chunk + got);
    const bool err = ferror(fp) != 0;
    fclose(fp);
    if (err) return fail(MST_E_IO, "read failed: " + path);
    if (buf.size() < 8 + sizeof(uint64_t) || memcmp(buf.data(), MESH_MAGIC, 8) != 0)
        return fail(MST_E_IO, path + ": not a mesh cache of this format");
    Fnv h;
    h.add(buf.data(), buf.size() - sizeof(uint64_t));
    uint64_t stored;
    memcpy(&stored, buf.data() + buf.size() - sizeof(uint64_t), sizeof stored);
    if (stored != h.h) return fail(MST_E_IO, path + ": checksum mismatch (damaged file)");
    Reader r{buf.data() + 8, buf.data() + buf.size() - sizeof(uint64_t)};
    if (r.val<uint32_t>() != 1) return fail(MST_E_IO, path + ": unknown version");
    mst_mesh t;
    t.dim = r.val<int32_t>();
    t.fnn = r.val<int32_t>();
    t.n_nodes = r.val<int64_t>();
    t.n_cells = r.val<int64_t>();
    t.n_faces = r.val<int64_t>();
    t.n_int = r.val<int64_t>();
    const int64_t nz = r.val<int64_t>();
    if (!r.ok || (t.dim != 2 && t.dim != 3) || t.fnn < 2 || t.fnn > 4 || t.n_nodes < 0 || t.n_faces < 0 || nz < 0)
        return fail(MST_E_IO, path + ": bad header");
    r.vec(t.xyz, t.n_nodes * t.dim);
    r.vec(t.fnodes, t.n_faces * t.fnn);
    r.vec(t.c0, t.n_faces);
    r.vec(t.c1, t.n_faces);
    r.vec(t.ftype, t.n_faces);
    r.vec(t.partner, t.n_faces);
    for (int64_t i = 0; i < nz && r.ok; i++) {
        mst_zone z;
        z.id = r.val<int32_t>();
        z.first = r.val<int32_t>();
        z.last = r.val<int32_t>();
        z.type = r.val<int32_t>();
        const int32_t len = r.val<int32_t>();
        if (len < 0 || len > (int32_t)(r.end - r.p)) {
            r.ok = false;
            break;
        }
        z.name.assign((const char *)r.p, (size_t)len);
        r.p += len;
        t.zones.push_back(z);
    }
    if (!r.ok || r.p != r.end) return fail(MST_E_IO, path + ": truncated or inconsistent");
    int rc = validate(t);
    if (rc) return rc;
    m = std::move(t);
    return MST_OK;
}

}  // namespace mst
