// Unit test of the host Newton control (include/dumux_amd/newtonsolver.hh) with the scalar
// mock assembler / linear solver of the reference's own test
// (test/nonlinear/newton/test_newton.cc:31-111): solve x^2 - 5 = 0 from x = 0.1 to 1e-13,
// plus the time-step retry path: a linear solver that fails twice forces two dt halvings.
#include "../../include/dumux_amd/newtonsolver.hh"

#include <cmath>
#include <cstdio>
#include <memory>

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
    std::printf("ok\n");
    return 0;
}
