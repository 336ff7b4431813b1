// rhosolver_driver.cpp -- the reference's time loop (Time::goNextTimeStep, time/Time.cpp:54-80) over
// mstcfd::RhoSolverGPU: per step setDT, solve, the max |New - Old| / Old monitor (:69-76),
// updateNewToOld.  Workload: the Sod strip of Time::initialization (:13-36) on NX x NY x 2 jittered
// triangles with walls (default 200 x 4), first-order Roe, forward Euler with a fixed DT (the
// reference's own setup).
// Usage: rhosolver_driver OUT NSTEPS DT [host|device [NX NY]] -- the monitor over the host mirrors of
// the states (the reference's loop: the new state crosses PCIe every step) or reduced on the device;
// writes the final state (file order, nv doubles per cell) to OUT and the loop's time per step.
// Exit status 3 with "ERR <message>" when the solver cannot be built (no HIP device).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>

#include "mstcfd.hpp"

int main(int argc, char **argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s OUT NSTEPS DT\n", argv[0]);
        return 2;
    }
    const char *out = argv[1];
    const int nsteps = std::atoi(argv[2]);
    const double dt = std::atof(argv[3]);
    const bool device_monitor = argc > 4 && std::strcmp(argv[4], "device") == 0;
    const int nx = argc > 6 ? std::atoi(argv[5]) : 200, ny = argc > 6 ? std::atoi(argv[6]) : 4;
    try {
        const int32_t walls[4] = {MST_BC_WALL, MST_BC_WALL, MST_BC_WALL, MST_BC_WALL};
        mstcfd::Mesh mesh = mstcfd::Mesh::genTri(nx, ny, 0.0, 1.0, 0.0, 0.005 * ny, 0.1, 12345, walls);
        mst_config cfg = {};
        cfg.scheme = MST_SCHEME_ROE;
        cfg.order = 1;
        cfg.limiter = MST_LIM_NONE;
        cfg.rk = 1;
        cfg.compat = MST_COMPAT_REFERENCE;
        cfg.precision = MST_PREC_FP64;
        cfg.gamma = 1.4;
        cfg.eor = 1e-10;
        cfg.entropy_delta = 0.125;
        cfg.cfl = 0;  // DT comes from setDT
        const double q0[4] = {1.0, 0.0, 0.0, 2.5};
        for (int k = 0; k < 4; k++) cfg.q_inlet[k] = q0[k];
        const std::vector<double> Q0 = mesh.initState(MST_INIT_REFERENCE, {q0[0], q0[1], q0[2], q0[3]});
        mstcfd::RhoSolverGPU s(mesh, cfg, Q0);
        const int64_t n = mesh.getNumOfCells();
        double mon[4] = {0, 0, 0, 0};
        s.getOldValue();
        const auto t0 = std::chrono::steady_clock::now();
        for (int it = 0; it < nsteps; it++) {
            s.setDT(dt);
            s.solve();
            if (device_monitor) {
                const std::vector<double> r = s.monitorMaxRel();
                for (int k = 0; k < 4; k++) mon[k] = r[k];
            } else {
                const double *qn = s.getNewValue(), *qo = s.getOldValue();
                for (int k = 0; k < 4; k++) mon[k] = 0;
                for (int64_t i = 0; i < n; i++)
                    for (int k = 0; k < 4; k++) {
                        const double v = std::fabs(qn[i * 4 + k] - qo[i * 4 + k]) / qo[i * 4 + k];
                        if (mon[k] < v) mon[k] = v;
                    }
            }
            s.updateNewToOld();
        }
        const double ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / nsteps;
        FILE *fp = std::fopen(out, "wb");
        if (!fp || std::fwrite(s.getOldValue(), sizeof(double), (size_t)n * 4, fp) != (size_t)n * 4) {
            std::printf("ERR cannot write %s\n", out);
            if (fp) std::fclose(fp);
            return 1;
        }
        std::fclose(fp);
        std::printf("monitor %.17e %.17e %.17e %.17e\nms_per_step %.6f (%s monitor)\nok %lld cells %d steps\n",
                    mon[0], mon[1], mon[2], mon[3], ms, device_monitor ? "device" : "host", (long long)n, nsteps);
        return 0;
    } catch (const std::exception &e) {
        std::printf("ERR %s\n", e.what());
        return 3;
    }
}
