// Unit test of the host Newton control (include/dumux_amd/newtonsolver.hh) with the scalar
// mock assembler / linear solver of the reference's own test
// (test/nonlinear/newton/test_newton.cc:31-111): solve x^2 - 5 = 0 from x = 0.1 to 1e-13,
// plus the time-step retry path: a linear solver that fails twice forces two dt halvings.
#include "../../include/dumux_amd/newtonsolver.hh"

#include <cmath>
#include <cstdio>
#include <memory>
#include <vector>

struct MockScalarAssembler {
    using SolutionVector = double;
    double res = 0, jac = 0, prev = 0.1;
    double dt = 1.0;
    void assembleResidual(const double& x) { res = x * x - 5.0; }
    void assembleJacobianAndResidual(const double& x) { assembleResidual(x); jac = 2.0 * x; }
    double& jacobian() { return jac; }
    double& residual() { return res; }
    double& update() { return du; }
    const double& prevSol() const { return prev; }
    void resetTimeStep(const double&) {}
    void updateGridVariables(const double&) {}
    bool isStationaryProblem() const { return false; }
    void setTimeStepSize(double v) { dt = v; }
    double du = 0;
};

struct MockScalarLinearSolver {
    int failures = 0;
    void setResidualReduction(double) {}
    bool solve(const double& A, double& x, const double& b)
    {
        if (failures > 0) { --failures; return false; }
        x = b / A;
        return true;
    }
};

struct ScalarBackend {
    static double clone(const double& u) { return u; }
    static void copy(double& dst, const double& src) { dst = src; }
    static double update(double& u, const double& uLast, const double& du)
    {
        u = uLast;
        u += -1.0 * du;
        return std::abs(uLast - u) / std::max(1.0, std::abs(uLast + u) * 0.5);
    }
};

// the scalar mock with a partial reassembler: records the hooks the Newton loop calls
// (newtonsolver.hh:415-419 resetColors, 565-603 the threshold, 1102-1108 the assembly)
struct MockPartialAssembler : MockScalarAssembler {
    int resets = 0, partialAssemblies = 0, fullAssemblies = 0;
    std::vector<double> thresholds;
    void resetPartialReassembly() { ++resets; }
    long long updatePartialReassembly(double t) { thresholds.push_back(t); return 0; }
    void assembleJacobianAndResidualPartial(const double& x) { ++partialAssemblies; MockScalarAssembler::assembleJacobianAndResidual(x); }
    void assembleJacobianAndResidual(const double& x) { ++fullAssemblies; MockScalarAssembler::assembleJacobianAndResidual(x); }
    std::size_t numDofs() const { return 1; }
};

int main()
{
    using Solver = Dumux::Gpu::NewtonSolver<MockScalarAssembler, MockScalarLinearSolver, ScalarBackend>;
    auto assembler = std::make_shared<MockScalarAssembler>();
    auto ls = std::make_shared<MockScalarLinearSolver>();
    Dumux::Gpu::NewtonParameters p;
    p.maxRelativeShift = 1e-13;
    p.maxSteps = 30;
    Solver solver(assembler, ls, p);
    double x = 0.1;
    solver.solve(x);
    const double err = std::abs(x - std::sqrt(5.0)) / std::sqrt(5.0);
    std::printf("x = %.15f err = %.3e steps = %d\n", x, err, solver.numSteps());
    if (!(err <= 1e-13)) return 1;
    if (solver.numSteps() < 2) return 2;   // Newton.MinSteps = 2

    // retry path: two linear failures -> two dt halvings, then convergence
    auto a2 = std::make_shared<MockScalarAssembler>();
    auto l2 = std::make_shared<MockScalarLinearSolver>();
    l2->failures = 2;
    Solver s2(a2, l2, p);
    Dumux::Gpu::TimeLoopState tl;
    tl.dt = 8.0;
    double y = 0.1;
    s2.solve(y, tl);
    std::printf("retries = %d dt = %g\n", s2.dtRetries(), tl.dt);
    if (s2.dtRetries() != 2 || tl.dt != 2.0 || a2->dt != 2.0) return 3;

    // exhausted divisions -> NumericalProblem
    auto a3 = std::make_shared<MockScalarAssembler>();
    auto l3 = std::make_shared<MockScalarLinearSolver>();
    l3->failures = 100;
    Dumux::Gpu::NewtonParameters p3 = p;
    p3.maxTimeStepDivisions = 3;
    Solver s3(a3, l3, p3);
    try {
        double z = 0.1;
        s3.solve(z, tl);
        return 4;
    } catch (const Dumux::Gpu::NumericalProblem&) {
    }
    // partial reassembly hooks: one reset per solve, the partial assembly in every iteration, and
    // after every update eps = max(minEps, min(maxEps, omega * shift)) with the defaults
    // minEps = 1e-1 * shiftTol, maxEps = 1e2 * shiftTol, omega = 1e-3
    {
        auto a4 = std::make_shared<MockPartialAssembler>();
        auto l4 = std::make_shared<MockScalarLinearSolver>();
        Dumux::Gpu::NewtonParameters p4 = p;
        p4.maxRelativeShift = 1e-8;
        p4.enablePartialReassembly = true;
        using PSolver = Dumux::Gpu::NewtonSolver<MockPartialAssembler, MockScalarLinearSolver, ScalarBackend>;
        PSolver s4(a4, l4, p4);
        double w = 0.1;
        s4.solve(w);
        const int n = s4.numSteps();
        if (a4->resets != 1 || a4->partialAssemblies != n || a4->fullAssemblies != 0 || int(a4->thresholds.size()) != n)
            return 5;
        // replay the shifts of the same iteration (same mock, same arithmetic)
        double v = 0.1, vl = 0.1;
        for (int k = 0; k < n; ++k) {
            const double du = (v * v - 5.0) / (2.0 * v);
            vl = v;
            v = vl;
            v += -1.0 * du;
            const double shift = std::abs(vl - v) / std::max(1.0, std::abs(vl + v) * 0.5);
            const double want = std::max(1e-9, std::min(1e-6, shift * 1e-3));
            if (a4->thresholds[k] != want) return 6;
        }
        if (!(a4->thresholds.front() == 1e-6 && a4->thresholds.back() == 1e-9)) return 7;   // both clamps reached
        if (s4.reassembledCells() != n) return 8;    // the mock reports no green cell
        // without the reassembler hooks the solver refuses to enable it
        Dumux::Gpu::NewtonParameters p5 = p;
        p5.enablePartialReassembly = true;
        Solver s5(std::make_shared<MockScalarAssembler>(), std::make_shared<MockScalarLinearSolver>(), p5);
        try {
            double z = 0.1;
            s5.solve(z);
            return 9;
        } catch (const std::logic_error&) {
        }
    }
    std::printf("ok\n");
    return 0;
}
