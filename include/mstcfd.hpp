/* mstcfd.hpp -- C++ host classes over the C ABI of mstcfd.h, with the reference solver plugin's method
 * set, so a driver reads like Time::goNextTimeStep (time/Time.cpp:54-80):
 *
 *     mstcfd::RhoSolverGPU s(mesh, cfg, Q0);
 *     s.setDT(dt); s.solve();                        // RhoSolver::setDT / solve (RhoSolver.cpp:33-68)
 *     monitor(s.getNewValue(), s.getOldValue());     // getNewValue / getOldValue (:504-509)
 *     s.updateNewToOld();                            // updateNewToOld (:513-517)
 *
 * Header only, C++17.  Errors throw std::runtime_error with the library's message (the reference has
 * no error returns); there is no CPU fallback: constructing a solver without a HIP device throws.
 * getOldValue / getNewValue return host mirrors in the reference's AllData layout (cell-major, nv
 * doubles per cell, file order), refreshed by solve() and updateNewToOld() as the reference's
 * arrays are.
 */
#pragma once
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mstcfd.h"

namespace mstcfd {

inline void check(int rc) {
    if (rc) {
        char buf[1024];
        mst_last_error(buf, sizeof buf);
        throw std::runtime_error(std::string("mstcfd: ") + buf);
    }
}

// the host mesh (MshBlock, mesh/MshBlock.h:19-76), file order
class Mesh {
public:
    static Mesh readMsh(const std::string &path) {  // MshBlock::readMsh (mesh/MshBlock.cpp:5-18)
        mst_mesh *m = nullptr;
        check(mst_mesh_read_msh(path.c_str(), &m));
        return Mesh(m);
    }
    static Mesh genTri(int nx, int ny, double x0, double x1, double y0, double y1, double jitter, uint64_t seed,
                       const int32_t bc[4]) {
        mst_mesh *m = nullptr;
        check(mst_mesh_gen_tri(nx, ny, x0, x1, y0, y1, jitter, seed, bc, &m));
        return Mesh(m);
    }
    Mesh(Mesh &&o) noexcept : m_(std::exchange(o.m_, nullptr)) {}
    Mesh &operator=(Mesh &&o) noexcept {
        std::swap(m_, o.m_);
        return *this;
    }
    Mesh(const Mesh &) = delete;
    Mesh &operator=(const Mesh &) = delete;
    ~Mesh() {
        if (m_) mst_mesh_destroy(m_);
    }
    const mst_mesh *get() const { return m_; }
    int getDim() const { return (int)count(0); }
    int64_t getNumOfCells() const { return count(3); }
    int64_t getNumOfFaces() const { return count(4); }
    // Time::initialization (time/Time.cpp:13-53) and the other initial fields of mst_mesh_init_state
    std::vector<double> initState(int kind, const std::vector<double> &params) const {
        std::vector<double> Q((size_t)getNumOfCells() * (getDim() + 2));
        check(mst_mesh_init_state(m_, kind, params.data(), (int)params.size(), Q.data(), getNumOfCells()));
        return Q;
    }

private:
    explicit Mesh(mst_mesh *m) : m_(m) {}
    int64_t count(int k) const {
        int64_t c[7];
        check(mst_mesh_counts(m_, c));
        return c[k];
    }
    mst_mesh *m_ = nullptr;
};

// the device solver with RhoSolver's method set (rhoSolver/RhoSolver.h:14-24); one context for the
// whole run, where the reference constructs its solver every step (time/Time.cpp:58).  The states live
// in HBM; the host arrays getOldValue / getNewValue return are mirrors filled only when they are read
// after the device state changed (a time loop that never reads them moves no state over PCIe).
class RhoSolverGPU {
public:
    RhoSolverGPU(const Mesh &mesh, const mst_config &cfg, const std::vector<double> &initial)
        : n_(mesh.getNumOfCells()), nv_(mesh.getDim() + 2) {
        check(mst_ctx_create(mesh.get(), &cfg, &ctx_));
        old_ = initial;
        new_ = initial;
        if ((int64_t)old_.size() != n_ * nv_) {
            mst_ctx_destroy(ctx_);
            throw std::runtime_error("mstcfd: initial state size does not match the mesh");
        }
        int rc = mst_set_state(ctx_, old_.data(), n_);
        if (rc) {
            mst_ctx_destroy(ctx_);
            check(rc);
        }
    }
    RhoSolverGPU(const RhoSolverGPU &) = delete;
    RhoSolverGPU &operator=(const RhoSolverGPU &) = delete;
    ~RhoSolverGPU() {
        if (ctx_) mst_ctx_destroy(ctx_);
    }

    void setDT(double dt) { check(mst_set_dt(ctx_, dt)); }  // RhoSolver.cpp:33-35
    void solve() {                                           // RhoSolver.cpp:37-68
        check(mst_solve(ctx_));
        new_is_old_ = false;
        new_stale_ = true;
    }
    const double *getOldValue() {  // RhoSolver.cpp:504-506
        if (old_stale_) {
            check(mst_get_state(ctx_, old_.data(), n_));
            old_stale_ = false;
        }
        return old_.data();
    }
    const double *getNewValue() {  // RhoSolver.cpp:507-509
        if (new_is_old_) return getOldValue();
        if (new_stale_) {
            check(mst_get_new_state(ctx_, new_.data(), n_));
            new_stale_ = false;
        }
        return new_.data();
    }
    void updateNewToOld() {  // RhoSolver.cpp:513-517: Old := New, and New keeps that value
        check(mst_update_new_to_old(ctx_));
        if (!new_is_old_) {
            if (!new_stale_) {
                old_.swap(new_);
                old_stale_ = false;
            } else {
                old_stale_ = true;
            }
        }
        new_is_old_ = true;
    }
    // the monitor of Time::goNextTimeStep (time/Time.cpp:69-76), max |New - Old| / Old per field, reduced
    // on the device: the time loop's per-step check without reading the states back
    std::vector<double> monitorMaxRel() {
        std::vector<double> r(nv_);
        check(mst_monitor_maxrel(ctx_, r.data()));
        return r;
    }

    // device-resident stepping (SSP-RK stages and the CFL time step of the configuration)
    void step(int n) {
        check(mst_step(ctx_, n));
        old_stale_ = new_is_old_ = true;
    }
    // the implicit steady driver (LU-SGS) to tol x the first iteration's norms
    int runToSteady(double tol, int maxIterations, int checkEvery, double cfl = 1e4, double omega = 1.0) {
        check(mst_set_steady_driver(ctx_, MST_STEADY_LUSGS, cfl, omega));
        std::vector<double> ref(5, 0.0);
        int steps = 0;
        double drop = 0;
        check(mst_run_to_steady(ctx_, tol, maxIterations, checkEvery, ref.data(), &steps, &drop));
        old_stale_ = new_is_old_ = true;
        return steps;
    }
    std::vector<double> residualNorms() const {
        std::vector<double> r(nv_);
        check(mst_residual_norms(ctx_, r.data()));
        return r;
    }
    mst_ctx *handle() const { return ctx_; }

private:
    mst_ctx *ctx_ = nullptr;
    int64_t n_;
    int nv_;
    std::vector<double> old_, new_;
    bool old_stale_ = false, new_stale_ = false;
    bool new_is_old_ = true;  // New equals Old (after updateNewToOld, step or the initial state)
};

}  // namespace mstcfd
