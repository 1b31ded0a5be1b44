// tf_driver.cpp -- minimal C++ host driver over tf_nr.hpp: reads a reference .mesh file, gives
// every prism the same neo-Hookean material, fixes basal z, grows uniformly in x and y and runs
// the implicit step the way Simulation::updateStepNR drives NewtonRaphsonSolver.
//   usage: tf_driver mesh_file [steps] [E] [nu] [growth_per_hour] [positions_out]
// positions_out receives the final node positions, one "x y z" line per node (%.17g).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "../../include/tf_nr.hpp"

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s mesh_file [steps] [E] [nu] [growth_per_hour] [positions_out]\n", argv[0]);
        return 2;
    }
    const int steps = argc > 2 ? std::atoi(argv[2]) : 3;
    const double E = argc > 3 ? std::atof(argv[3]) : 1000.0;
    const double nu = argc > 4 ? std::atof(argv[4]) : 0.3;
    const double g_hr = argc > 5 ? std::atof(argv[5]) : 0.01;
    std::ifstream f(argv[1]);
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", argv[1]);
        return 2;
    }
    int n = 0;
    f >> n;
    std::vector<double> pos(3 * n);
    std::vector<int> placement(n);
    std::vector<uint8_t> fixed(3 * n, 0);
    for (int i = 0; i < n; ++i) {
        int tp, tt, circ;
        f >> pos[3 * i] >> pos[3 * i + 1] >> pos[3 * i + 2] >> tp >> tt >> circ;
        placement[i] = tp;
        if (tp == 0) fixed[3 * i + 2] = 1;  // basal z fixed
    }
    int ne = 0;
    f >> ne;
    std::vector<int32_t> conn(6 * ne);
    std::vector<double> ref(18 * ne);
    for (int e = 0; e < ne; ++e) {
        int shape;
        f >> shape;
        for (int a = 0; a < 6; ++a) f >> conn[6 * e + a];
        for (int k = 0; k < 18; ++k) f >> ref[18 * e + k];
    }
    // pin the in-plane rigid modes at the first basal node and one more basal node
    int pinned = 0;
    for (int i = 0; i < n && pinned < 2; ++i)
        if (placement[i] == 0) {
            fixed[3 * i + 1] = 1;
            if (pinned == 0) fixed[3 * i] = 1;
            ++pinned;
        }
    const double lam = E * nu / (1 + nu) / (1 - 2 * nu), mu = E / 2 / (1 + nu);
    std::vector<double> L(ne, lam), M(ne, mu), rate(3 * ne, 0.0);
    for (int e = 0; e < ne; ++e) rate[3 * e] = rate[3 * e + 1] = g_hr / 3600.0;
    tf_mesh_desc d{};
    d.n_nodes = n;
    d.n_elements = ne;
    d.node_pos = pos.data();
    d.conn = conn.data();
    d.ref_pos = ref.data();
    d.lambda = L.data();
    d.mu = M.data();
    d.fixed = fixed.data();
    d.growth_rate = rate.data();
    d.n_segments = 1;
    try {
        tf::NewtonRaphsonSolver s(d);
        const double dt = 60.0;
        for (int step = 0; step < steps; ++step) {
            s.growShapeByFg(dt);
            s.updateNodeViscositySurfaces();
            s.constructUnMatrix();
            bool converged = false;
            int it = 0;
            while (!converged) {
                s.calculateForcesAndJacobianMatrixNR(dt);
                s.solveForDeltaU();
                converged = s.checkConvergenceViaDeltaU();
                std::printf("step %d iteration %d |dU| %.3e krylov %d\n", step, it, s.lastDeltaUNorm,
                            s.lastKrylovIterations);
                if (!converged && ++it > 20) break;
            }
        }
        std::vector<double> x = s.positions();
        std::printf("node 0 at (%.9f, %.9f, %.9f)\n", x[0], x[1], x[2]);
        if (argc > 6) {
            FILE* out = std::fopen(argv[6], "w");
            if (!out) {
                std::fprintf(stderr, "cannot write %s\n", argv[6]);
                return 2;
            }
            for (int i = 0; i < n; ++i) std::fprintf(out, "%.17g %.17g %.17g\n", x[3 * i], x[3 * i + 1], x[3 * i + 2]);
            std::fclose(out);
        }
    } catch (const tf::Error& ex) {
        std::fprintf(stderr, "tf error %d: %s\n", ex.code, ex.what());
        return 1;
    }
    return 0;
}
