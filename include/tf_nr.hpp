// tf_nr.hpp -- header-only C++ facade over the C ABI in tf_nr.h.
//
// tf::NewtonRaphsonSolver keeps the reference's NewtonRaphsonSolver method names
// (TissueFolding/TissueFolding/SourceCode/NewtonRaphsonSolver.h:26-88) so that the driver loop of
// Simulation::updateStepNR (Simulation.cpp:6449-6590) reads the same against the GPU path:
//
//     solver.constructUnMatrix();                 // un <- x, uk <- un
//     while (!converged) {
//         solver.calculateForcesAndJacobianMatrixNR(dt);   // forces, Ke/Kv, scatter, BCs
//         solver.solveForDeltaU();
//         converged = solver.checkConvergenceViaDeltaU();  // and updateUkInIteration
//     }
//
// Element-side entry points of ShapeBase/Prism (calculateForces, calculateImplicitKElastic,
// calculateImplicitKViscous) are folded into calculateForcesAndJacobianMatrixNR, which runs them
// for every prism on the GPU; elementDump() returns one prism's arrays in the reference layout.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "tf_nr.h"

namespace tf {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code(code) {}
    int code;
};

class NewtonRaphsonSolver {
public:
    explicit NewtonRaphsonSolver(const tf_mesh_desc& desc)
    {
        check(tf_create(&desc, &h_), nullptr);
        int32_t n = 0, e = 0, nb = 0, s = 0;
        tf_get_sizes(h_, &n, &e, &nb, &s);
        nNodes_ = n;
        nElements_ = e;
    }
    ~NewtonRaphsonSolver() { tf_destroy(h_); }
    NewtonRaphsonSolver(const NewtonRaphsonSolver&) = delete;
    NewtonRaphsonSolver& operator=(const NewtonRaphsonSolver&) = delete;

    tf_solver_opts options() const
    {
        tf_solver_opts o;
        tf_get_options(h_, &o);
        return o;
    }
    void setOptions(const tf_solver_opts& o) { check(tf_set_options(h_, &o), h_); }

    // ---- per step (Simulation::runOneStep) ----
    int growShapeByFg(double dt)
    {
        int32_t n = 0;
        check(tf_grow(h_, dt, &n), h_);
        return n;
    }
    void updateNodeViscositySurfaces() { check(tf_update_viscosity_surfaces(h_), h_); }

    // NewtonRaphsonSolver::slaveMasterList (each entry {slave DoF, master DoF}); the binding is
    // applied by every later assembly (calculateBoundKWithSlavesMasterDoF) and solve
    // (equateSlaveDisplacementsToMasters).
    void setSlaveMasterList(const std::vector<std::vector<int>>& slaveMasterList)
    {
        std::vector<int32_t> p;
        for (const auto& sm : slaveMasterList) {
            if (sm.size() != 2) throw Error(TF_INVALID, "slave/master entries hold two DoFs");
            p.push_back(sm[0]);
            p.push_back(sm[1]);
        }
        check(tf_set_bound_dofs(h_, (int32_t)slaveMasterList.size(), p.empty() ? nullptr : p.data()), h_);
    }
    // Simulation::AblatedNodes, used by checkJacobianForAblatedNodes.
    void setAblatedNodes(const std::vector<int>& nodes)
    {
        std::vector<int32_t> n(nodes.begin(), nodes.end());
        check(tf_set_ablated_nodes(h_, (int32_t)n.size(), n.empty() ? nullptr : n.data()), h_);
    }

    // Simulation::updateNodeMasses + detectPacingNodes on the current positions; the pairs act on
    // every later assembly (calculatePackingForcesImplicit3D, calculatePackingJacobian3D).
    int detectPacingNodes()
    {
        int32_t n = 0;
        check(tf_detect_packing(h_, &n), h_);
        return n;
    }
    // pacingNodeCouples0/1 with initialWeightPointx/y/z and packingThreshold (one per tissue),
    // e.g. from the reference's own detection or a Save_Packing file.
    void setPackingPairs(const std::vector<int>& couples0, const std::vector<int>& couples1,
                         const std::vector<double>& wx, const std::vector<double>& wy, const std::vector<double>& wz,
                         const std::vector<double>& packingThreshold)
    {
        const size_t n = couples0.size();
        if (couples1.size() != n || wx.size() != n || wy.size() != n || wz.size() != n)
            throw Error(TF_INVALID, "packing couples and weights differ in length");
        std::vector<int32_t> p(2 * n);
        std::vector<double> w(3 * n);
        for (size_t k = 0; k < n; ++k) {
            p[2 * k] = couples0[k];
            p[2 * k + 1] = couples1[k];
            w[3 * k] = wx[k];
            w[3 * k + 1] = wy[k];
            w[3 * k + 2] = wz[k];
        }
        check(tf_set_packing_pairs(h_, (int32_t)n, n ? p.data() : nullptr, n ? w.data() : nullptr,
                                   packingThreshold.data()),
              h_);
    }
    // Node::mass of every node from the current Fg
    std::vector<double> nodeMasses()
    {
        std::vector<double> m(nNodes_);
        check(tf_get_node_masses(h_, m.data()), h_);
        return m;
    }

    // ---- NewtonRaphsonSolver call sequence ----
    void constructUnMatrix() { check(tf_begin_step(h_), h_); }
    void initialteUkMatrix() {}  // folded into constructUnMatrix (uk <- un)
    int calculateForcesAndJacobianMatrixNR(double dt)
    {
        const int rc = tf_assemble(h_, dt);
        if (rc != TF_OK && rc != TF_NEG_JACOBIAN) check(rc, h_);
        return rc;
    }
    int solveForDeltaU(std::vector<double>* deltaU = nullptr)
    {
        if (deltaU) deltaU->resize(3 * (size_t)nNodes_);
        int32_t it = 0;
        double rr = 0.0;
        const int rc = tf_solve(h_, deltaU ? deltaU->data() : nullptr, &it, &rr);
        if (rc != TF_OK && rc != TF_KRYLOV_STALL) check(rc, h_);
        lastKrylovIterations = it;
        return rc;
    }
    bool checkConvergenceViaDeltaU()
    {
        int32_t conv = 0;
        check(tf_update_uk(h_, &lastDeltaUNorm, &conv), h_);
        return conv != 0;
    }
    std::vector<double> gSum()
    {
        std::vector<double> g(3 * (size_t)nNodes_);
        check(tf_get_residual(h_, g.data()), h_);
        return g;
    }
    // whole updateStepNR on the device
    tf_step_stats updateStepNR(double dt)
    {
        tf_step_stats st;
        const int rc = tf_nr_step(h_, dt, &st);
        if (rc != TF_OK && rc != TF_NOT_CONVERGED) check(rc, h_);
        return st;
    }
    // growth + surfaces + updateStepNR
    tf_step_stats runOneStep(double dt)
    {
        tf_step_stats st;
        const int rc = tf_timestep(h_, dt, &st);
        if (rc != TF_OK && rc != TF_NOT_CONVERGED) check(rc, h_);
        return st;
    }

    // ---- state ----
    std::vector<double> positions()
    {
        std::vector<double> x(3 * (size_t)nNodes_);
        check(tf_get_positions(h_, x.data()), h_);
        return x;
    }
    void setPositions(const std::vector<double>& x) { check(tf_set_positions(h_, x.data()), h_); }
    // ShapeBase::setFg for every element: Fg row-major 3x3 per element (InvFg is derived)
    void setFg(const std::vector<double>& Fg) { check(tf_set_fg(h_, Fg.data()), h_); }
    // ShapeBase::growthIncrement of every element (row-major 3x3), as Simulation::calculateGrowth
    // builds it from the growth functions; an empty vector returns to the growth rates
    void setGrowthIncrements(const std::vector<double>& dFg)
    {
        check(tf_set_growth_increments(h_, dFg.empty() ? nullptr : dFg.data()), h_);
    }
    void elementDump(int e, double dt, double fe[18], double fv[18], double Ke[324], double Kv[324])
    {
        check(tf_element_dump(h_, e, dt, fe, fv, Ke, Kv), h_);
    }

    int nNodes() const { return nNodes_; }
    int nElements() const { return nElements_; }
    tf_handle* handle() { return h_; }

    int lastKrylovIterations = 0;
    double lastDeltaUNorm = 0.0;

private:
    static void check(int rc, const tf_handle* h)
    {
        if (rc != TF_OK) throw Error(rc, tf_last_error(h));
    }
    tf_handle* h_ = nullptr;
    int nNodes_ = 0, nElements_ = 0;
};

}  // namespace tf
