// mesh.hpp -- host-side mesh of the MI355X residual path.
//
// mst_mesh is the reference's MshBlock (mesh/MshBlock.h:19-76) reduced to flat arrays in FILE order:
// faces with their nodes and two cells (c1 = -1 on boundary faces), a zone table (FacesInf,
// mesh/FacesInf.h) and, new, a periodic partner per face.  Everything here is integer work plus
// the FP64 node coordinates; it is computed once on the host.
#pragma once
#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "mstcfd.h"

struct mst_zone {
    int32_t id;
    int32_t first;  // first face (0-based)
    int32_t last;   // one past the last face
    int32_t type;   // Fluent BC type
    std::string name;
};

struct mst_mesh {
    int dim = 2;
    int fnn = 2;  // nodes per face (2 in 2D, 3 for triangle faces in 3D)
    int64_t n_nodes = 0, n_cells = 0, n_faces = 0;
    int64_t n_int = 0;  // numOfIntFaces as the reference reader defines it (MshBlock.cpp:223-226)
    std::vector<double> xyz;         // n_nodes * dim
    std::vector<int32_t> fnodes;     // n_faces * fnn
    std::vector<int32_t> c0, c1;     // n_faces
    std::vector<int32_t> ftype;      // n_faces
    std::vector<int32_t> partner;    // n_faces, periodic partner or -1
    std::vector<mst_zone> zones;
};

namespace mst {

// thread-local error text behind mst_last_error
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

// zone table from contiguous runs of ftype; n_int from the last interior run
void rebuild_zones(mst_mesh &m);
// structural checks (ids in range, every cell closed by >= dim+1 faces, partner symmetric)
int validate(const mst_mesh &m);

int read_msh(const std::string &path, mst_mesh &m);
int write_msh(const mst_mesh &m, const std::string &path);

int gen_tri(int nx, int ny, double x0, double x1, double y0, double y1, double jitter, uint64_t seed,
            const int32_t bc[4], mst_mesh &m);
int gen_tet(int nx, int ny, int nz, const double lo[3], const double hi[3], double jitter, uint64_t seed,
            const int32_t bc[6], mst_mesh &m);
int gen_ramp(int nx, int ny, double x0, double x1, double height, double theta_deg, double jitter, uint64_t seed,
             const int32_t bc[4], mst_mesh &m);
// cell centres (mean of face centres, mesh/Cell.cpp:6-14) in file order
int cell_centres(const mst_mesh &m, std::vector<double> &ccen);

// on-disk formats (output.cpp): Tecplot ASCII result file, binary checkpoint
int write_tecplot(const mst_mesh &m, const double *Q, double gamma, double cv, int wall_noslip, const std::string &title,
                  const std::string &path);
int write_checkpoint(const std::string &path, const mst_mesh &m, const double *Q, int64_t step, double time);
int read_checkpoint(const std::string &path, const mst_mesh &m, double *Q, int64_t *step, double *time);
int write_mesh_cache(const mst_mesh &m, const std::string &path);
int read_mesh_cache(const std::string &path, mst_mesh &m);

// std::mt19937_64 (its output sequence is fixed by the C++ standard) with a portable 53-bit
// uniform in [0,1), so generated meshes are identical on every host
struct Rng {
    std::mt19937_64 eng;
    explicit Rng(uint64_t seed) : eng(seed) {}
    double uniform() { return (double)(eng() >> 11) * (1.0 / 9007199254740992.0); }
    double sym() { return 2.0 * uniform() - 1.0; }  // [-1,1)
};

}  // namespace mst
