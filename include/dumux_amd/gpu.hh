// C++ adapters over the dumux_amd C ABI with the member functions DuMux's NewtonSolver calls:
//   Dumux::Gpu::FVAssembler   ~ Dumux::FVAssembler            (dumux/assembly/fvassembler.hh:101-662)
//   Dumux::Gpu::LinearSolver  ~ Dumux::ILUBiCGSTABIstlSolver  (dumux/linear/istlsolvers.hh:635-642)
//   Dumux::Gpu::DeviceBackend ~ the Backend ops of NewtonSolver (axpy + maxRelativeShift)
// Solution, residual, update and Jacobian stay in device memory; the vector "types" are
// handles to the context's buffers. Errors from the C ABI become exceptions:
// DMX_NUMERICAL_PROBLEM -> Dumux::Gpu::NumericalProblem, fatal codes -> std::runtime_error.
#pragma once
#include "../dumux_amd.h"
#include "newtonsolver.hh"

#include <memory>
#include <stdexcept>
#include <string>

namespace Dumux::Gpu {

inline void check(int rc, const dmx_context* ctx, const char* what)
{
    if (rc == DMX_OK) return;
    std::string msg = std::string(what) + ": " + (ctx ? dmx_last_error(ctx) : "");
    if (rc == DMX_NUMERICAL_PROBLEM) throw NumericalProblem(msg);
    throw std::runtime_error(msg);
}

struct DeviceVector {
    enum Kind { Solution, LastIterate, Previous, Residual, Update, Jacobian };
    dmx_context* ctx;
    Kind kind;
};

class FVAssembler {
public:
    using Scalar = double;
    using SolutionVector = DeviceVector;
    using ResidualType = DeviceVector;
    using JacobianMatrix = DeviceVector;

    // non-owning view of an existing context
    explicit FVAssembler(dmx_context* ctx) : ctx_(ctx) { bindHandles(); }
    // owning: FVAssembler(problem, gridGeometry, gridVariables[, timeLoop, prevSol]) equivalent
    FVAssembler(const dmx_problem_desc& desc, int device = 0)
    {
        check(dmx_create(&desc, device, &ctx_), nullptr, "dmx_create");
        owned_ = true;
        bindHandles();
    }
    ~FVAssembler()
    {
        if (owned_) dmx_destroy(ctx_);
    }
    FVAssembler(const FVAssembler&) = delete;
    FVAssembler& operator=(const FVAssembler&) = delete;

    dmx_context* context() const { return ctx_; }
    void setLinearSystem() {}  // pattern is built at construction (fvassembler.hh:294-302)
    std::size_t numDofs() const
    {
        int64_t n = 0, nnzb = 0;
        int32_t m = 0;
        dmx_num_dofs(ctx_, &n, &m, &nnzb);
        return std::size_t(n);
    }
    void assembleJacobianAndResidual(const SolutionVector&)
    {
        check(dmx_assemble_jacobian_and_residual(ctx_), ctx_, "assembleJacobianAndResidual");
    }
    void assembleResidual(const SolutionVector&) { check(dmx_assemble_residual(ctx_), ctx_, "assembleResidual"); }
    // the partial reassembler (dumux/assembly/partialreassembler.hh) kept on the device
    void resetPartialReassembly() { check(dmx_partial_reassembly_reset(ctx_), ctx_, "resetColors"); }
    long long updatePartialReassembly(double threshold)
    {
        int64_t green = 0;
        check(dmx_partial_reassembly_update(ctx_, threshold, &green), ctx_, "computeColors");
        return green;
    }
    void assembleJacobianAndResidualPartial(const SolutionVector&)
    {
        check(dmx_assemble_jacobian_and_residual_partial(ctx_), ctx_, "assembleJacobianAndResidual(partial)");
    }
    JacobianMatrix& jacobian() { return jac_; }
    ResidualType& residual() { return res_; }
    ResidualType& update() { return upd_; }
    const SolutionVector& prevSol() const { return prev_; }
    void setPreviousSolution(const double* uprevHost) { check(dmx_set_prev_solution(ctx_, uprevHost), ctx_, "setPreviousSolution"); }
    void resetTimeStep(const SolutionVector&) { check(dmx_reset_time_step(ctx_), ctx_, "resetTimeStep"); }
    void advanceTimeStep() { check(dmx_advance_time_step(ctx_), ctx_, "advanceTimeStep"); }
    void updateGridVariables(const SolutionVector&) {}  // volume variables are recomputed per assembly
    bool isStationaryProblem() const { return stationary_; }
    void setStationary(bool s) { stationary_ = s; }
    void setTimeStepSize(double dt) { check(dmx_set_time_step_size(ctx_, dt), ctx_, "setTimeStepSize"); }
    SolutionVector& solution() { return sol_; }

private:
    dmx_context* ctx_ = nullptr;
    bool owned_ = false;
    bool stationary_ = false;
    DeviceVector sol_{nullptr, DeviceVector::Solution};
    DeviceVector jac_{nullptr, DeviceVector::Jacobian};
    DeviceVector res_{nullptr, DeviceVector::Residual};
    DeviceVector upd_{nullptr, DeviceVector::Update};
    DeviceVector prev_{nullptr, DeviceVector::Previous};

public:
    void bindHandles()
    {
        sol_.ctx = jac_.ctx = res_.ctx = upd_.ctx = prev_.ctx = ctx_;
    }
};

class LinearSolver {
public:
    explicit LinearSolver(const dmx_linear_params& p) : params_(p) {}
    void setResidualReduction(double r) { params_.reduction = r; }
    void setMaxIter(int it) { params_.max_iter = it; }
    const std::string& name() const { return name_; }
    // solve(A, x, b) (istlsolvers.hh:273): A = jacobian, x = update, b = residual
    bool solve(DeviceVector& A, DeviceVector&, DeviceVector&)
    {
        dmx_solve_info info{};
        const int rc = dmx_linear_solve(A.ctx, &params_, &info);
        if (rc < 0) check(rc, A.ctx, "linear solve");
        iterations_ += info.iterations;
        refinements_ += info.refinements;
        return rc == DMX_OK && info.converged;
    }
    double norm(DeviceVector& b)
    {
        double n = 0.0;
        check(dmx_residual_norm(b.ctx, &n), b.ctx, "norm");
        return n;
    }
    int iterations() const { return iterations_; }
    int refinements() const { return refinements_; }

private:
    dmx_linear_params params_;
    std::string name_ = "BiCGSTAB (dumux_amd)";
    int iterations_ = 0, refinements_ = 0;
};

struct DeviceBackend {
    static DeviceVector clone(const DeviceVector& u) { return DeviceVector{u.ctx, DeviceVector::LastIterate}; }
    // uLastIter <- u is implicit: dmx_newton_update keeps the old iterate in a second buffer.
    // u <- prevSol is resetTimeStep.
    static void copy(DeviceVector& dst, const DeviceVector& src)
    {
        if (dst.kind == DeviceVector::Solution && src.kind == DeviceVector::Previous)
            check(dmx_reset_time_step(dst.ctx), dst.ctx, "reset");
    }
    static double update(DeviceVector& u, const DeviceVector&, const DeviceVector&)
    {
        double shift = 0.0;
        check(dmx_newton_update(u.ctx, &shift), u.ctx, "newtonUpdate");
        return shift;
    }
};

using DeviceNewtonSolver = NewtonSolver<FVAssembler, LinearSolver, DeviceBackend>;

} // namespace Dumux::Gpu
