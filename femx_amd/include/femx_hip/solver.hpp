// femx_hip/solver.hpp -- femx's linear-solve contract over the HIP backend
// (femx/linalg/LinearSolver.hpp:10-47) and a device BiCGStab that fills it.
//
// femx's device solver is ReSolve, which is not part of this path.
// BiCGStabSolver is a small deterministic replacement built on this
// library's kernels only -- CSR SpMV, the fixed-order dot product, axpby and
// a Jacobi scaling -- so that a reduced-functional gradient (forward solve,
// adjoint solve with the transposed Jacobian, parameter VJP) runs on the GPU.
// solveT solves with the explicit transpose, whose pattern and permutation
// are cached per source layout (MatrixHandler::transpose).  It is a
// correctness-grade solver for the gradient loop, not a tuned one.
#pragma once

#include <cmath>
#include <map>
#include <memory>

#include "femx_hip.hpp"

namespace femx::linalg
{

/** The linear solve contract (LinearSolver<Backend>, LinearSolver.hpp:10-47)
 * for the HIP backend. */
class LinearSolver
{
public:
  using Matrix  = DeviceCsrMatrix;
  using Vector  = DeviceVector;
  using Context = HipContext;

  virtual ~LinearSolver() = default;
  /** sol <- mat^-1 rhs; sol holds the initial guess on entry. */
  virtual void solve(const Matrix& mat, const Vector& rhs, Vector& sol, Context& ctx) = 0;
  /** sol <- mat^-T rhs; sol holds the initial guess on entry. */
  virtual void solveT(const Matrix& mat, const Vector& rhs, Vector& sol, Context& ctx) = 0;
};

/** Right-preconditioned (Jacobi) BiCGStab with every reduction in
 * fx_vec_dot's fixed order: repeated solves are bitwise identical.  Stops
 * when ||r|| <= tol ||b||; throws on breakdown or after max_iter steps. */
class BiCGStabSolver final : public LinearSolver
{
public:
  explicit BiCGStabSolver(Real tol = 1e-12, Index max_iter = 5000) : tol_(tol), max_iter_(max_iter) {}

  Index iterations() const noexcept { return iterations_; }
  Real  relResidual() const noexcept { return rel_; }

  void solve(const Matrix& A, const Vector& b, Vector& x, Context& ctx) override
  {
    require(A.rows() == A.cols() && b.size() == A.rows(), "BiCGStab needs a square system");
    const Index n = A.rows();
    if (x.size() != n) x.resize(n);
    HipMatrixHandler mh(ctx);
    HipVectorHandler vh(ctx);
    for (auto* v : {&d_, &r_, &rh_, &p_, &v_, &ph_, &s_, &sh_, &t_})
      if (v->size() != n) v->resize(n);
    if (dot_.size() != 1) dot_.resize(1);
    const auto& pos = diagPositions(A, ctx);
    vh.gather(view(A.vals()), view(pos), view(d_));
    auto dot = [&](const Vector& a, const Vector& c) {
      vh.dot(view(a), view(c), view(dot_));
      return dot_.toHost(ctx)[0];
    };
    auto div = [&](const Vector& a, Vector& z) {
      check(fx_vec_div(n, a.data(), d_.data(), z.data(), ctx.raw()));
    };
    mh.matvec(A, view(x), view(r_), -1.0, 0.0);
    vh.axpby(1.0, view(b), 1.0, view(r_));
    vh.copy(view(r_), view(rh_));
    vh.zero(view(p_));
    vh.zero(view(v_));
    const Real bnorm = std::sqrt(dot(b, b));
    iterations_      = 0;
    if (bnorm == 0.0)
    {
      vh.zero(view(x));
      rel_ = 0.0;
      return;
    }
    Real rho = 1.0, alpha = 1.0, omega = 1.0;
    Real rel = std::sqrt(dot(r_, r_)) / bnorm;
    while (rel > tol_)
    {
      require(iterations_ < max_iter_, "BiCGStab did not converge");
      ++iterations_;
      const Real rho_new = dot(rh_, r_);
      require(rho_new != 0.0 && omega != 0.0, "BiCGStab broke down");
      const Real beta = (rho_new / rho) * (alpha / omega);
      vh.axpby(-omega, view(v_), 1.0, view(p_));
      vh.axpby(1.0, view(r_), beta, view(p_));
      div(p_, ph_);
      mh.matvec(A, view(ph_), view(v_));
      alpha = rho_new / dot(rh_, v_);
      vh.copy(view(r_), view(s_));
      vh.axpby(-alpha, view(v_), 1.0, view(s_));
      const Real snorm = std::sqrt(dot(s_, s_)) / bnorm;
      if (snorm <= tol_)
      {
        vh.axpby(alpha, view(ph_), 1.0, view(x));
        rel = snorm;
        break;
      }
      div(s_, sh_);
      mh.matvec(A, view(sh_), view(t_));
      const Real tt = dot(t_, t_);
      omega         = tt != 0.0 ? dot(t_, s_) / tt : 0.0;
      vh.axpby(alpha, view(ph_), 1.0, view(x));
      vh.axpby(omega, view(sh_), 1.0, view(x));
      vh.copy(view(s_), view(r_));
      vh.axpby(-omega, view(t_), 1.0, view(r_));
      rel = std::sqrt(dot(r_, r_)) / bnorm;
      rho = rho_new;
    }
    rel_ = rel;
  }

  void solveT(const Matrix& A, const Vector& b, Vector& x, Context& ctx) override
  {
    auto& at = transposes_[A.pattern().layoutId()];
    HipMatrixHandler(ctx).transpose(A, at);
    solve(at, b, x, ctx);
  }

private:
  /** CSR position of each row's diagonal, cached per layout. */
  const DeviceIndexVector& diagPositions(const Matrix& A, Context& ctx)
  {
    auto& d = diag_[A.pattern().layoutId()];
    if (d.size() != A.rows())
    {
      const auto&        h = A.pattern().host();
      std::vector<Index> pos(h.rows, -1);
      for (Index r = 0; r < h.rows; ++r)
        for (Index k = h.row_ptr[r]; k < h.row_ptr[r + 1]; ++k)
          if (h.col_ind[k] == r)
          {
            pos[r] = k;
            break;
          }
      for (Index p : pos) require(p >= 0, "BiCGStab's Jacobi scaling needs a diagonal entry in every row");
      d.assign(pos, ctx);
    }
    return d;
  }

  Real                                        tol_;
  Index                                       max_iter_;
  Index                                       iterations_ = 0;
  Real                                        rel_        = 0.0;
  Vector                                      d_, r_, rh_, p_, v_, ph_, s_, sh_, t_, dot_;
  std::map<std::uint64_t, DeviceIndexVector>  diag_;
  std::map<std::uint64_t, DeviceCsrMatrix>    transposes_;
};

} // namespace femx::linalg
