// prep.hpp -- one-time host preprocessing of a file-order mesh into the device layout.
#pragma once
#include <cstdint>
#include <vector>

#include "mesh.hpp"

namespace mst {

// face kinds as the kernels switch on them (Fluent types of rhoSolver/RhoSolver.cpp:97-230)
enum FaceKind : int32_t {
    KIND_INTERIOR = 0,
    KIND_WALL = 1,
    KIND_SYMMETRY = 2,
    KIND_INLET = 3,
    KIND_OUTLET = 4,
    KIND_PERIODIC = 5,
    KIND_NONE = 6,   // unknown Fluent type: the reference skips it, the flux stays zero
    KIND_MERGED = 7  // second face of a periodic pair: its flux is minus the first face's
};

struct Prepared {
    int dim = 2, nv = 4, maxf = 3, ncolours = 0;
    int64_t nc = 0, nf = 0;
    int64_t n_interior = 0;           // new-order faces [0, n_interior) have two cells
    int64_t pstart = 0, pcount = 0;   // periodic faces are the contiguous range [pstart, pstart+pcount)
    int64_t quirk_face = -1;          // new index of file face N_int-1 (RhoSolver.cpp:437-440), -1 if none
    int64_t bad_cells = 0;            // cells whose reference area formula gives <= 0 or NaN (degenerate)
    int64_t first_bad_cell = -1;

    // geometry in FILE order (what the reference stores on Face/Cell)
    std::vector<double> fcen, S, eta;  // nf*dim, nf*dim, nf
    std::vector<int32_t> dac;          // directAndCells: -1 if S points into c0 (MshBlock.cpp:282-283)
    std::vector<double> ccen, vol;     // nc*dim, nc

    // integer preprocessing
    std::vector<int32_t> cell_new_of_old, cell_old_of_new;
    std::vector<int32_t> face_new_of_old, face_old_of_new;
    std::vector<int32_t> colour_new;   // per new face

    // device layout, NEW order
    std::vector<int32_t> fc0, fc1;      // fc1: neighbour (interior), partner face's cell (periodic), -1
    std::vector<int32_t> fkind;         // FaceKind
    std::vector<uint8_t> ffile_branch2; // 1 if the file face had S pointing out of c0 (MshBlock.cpp:294)
    std::vector<double> fn;             // nf*dim: outward normal of c0, directAndCells * S
    std::vector<double> fxf;            // nf*dim: face centre
    std::vector<double> feta;           // nf: eta0 (Face.cpp:62-69)
    std::vector<double> pxf;            // pcount*dim: centre of the partner face (periodic)
    std::vector<int32_t> fmerged_into;  // nf: for KIND_MERGED faces the new index of their pair's first face
    std::vector<double> cvol, cxc, cshalf;  // nc, nc*dim, nc*dim (half sums of |S_d|)
    std::vector<int32_t> ell;           // maxf*nc, slot-major: (face << 1) | (cell is c1); -1 padding
};

int prepare(const mst_mesh &m, Prepared &P);

// process-wide cell order of prepare (MST_ORDER_HILBERT or MST_ORDER_RCM)
int cell_order();
void set_cell_order(int order);

// per new face: bit d set when c0 (or the interior-side state) is the Riemann left state along d
void lr_masks(const Prepared &P, int compat, std::vector<int32_t> &mask);

// Incidence entry of slot s of cell c (the gradient kernel's layout): code = (neighbour << 1) | is_c1,
// -1 for a boundary face, -2 for an empty slot; n = the cell's outward normal; dx = x_f - x_c (the
// partner face's centre on the c1 side of a periodic pair).  Returns the face (or -1).
int64_t incidence(const Prepared &P, int s, int64_t c, int32_t &code, double *n, double *dx);

// colouring of the cell graph for the LU-SGS sweeps (new cell order): no two cells sharing a face
// alike; two colours when the graph is bipartite, else greedy + Kempe-chain reduction.  Returns the
// number of colours.
int lusgs_colours(const Prepared &P, std::vector<int32_t> &col);

}  // namespace mst

// the opaque handle of the C ABI
struct mst_prep {
    mst::Prepared P;
};
