// femx_hip/time.hpp -- femx's time-residual contract over the HIP backend,
// its Navier-Stokes model residual, Dirichlet control and the constrained
// time residual.  Header-only, over femx_hip.hpp:
//
//   state::TimeDims, TimeHistoryView,     femx/state/TimeResidual.hpp:13-88
//     TimeContext
//   state::TimeResidual                   femx/state/TimeResidual.hpp:170-239
//   model::ns::NavierResidual             NavierStokesModel::Residual
//                                         (femx/model/ns/NavierStokesModel.cpp:446-545)
//   fem::DirichletControl                 femx/fem/DirichletControl.hpp:22-60,
//                                         DirichletControl.cpp:21-73
//   fem::LinearInterpolation              femx/common/LinearInterpolation.hpp:12-40
//   fem::ControlMap, makeControlMap,      femx/fem/ControlMap.hpp:23-182,
//     copy, controlVals, controlJac,      ControlMap.cpp:36-342
//     addControlJacT
//   assembly::ConstrainedTimeResidual     femx/assembly/ConstrainedTimeResidual.hpp:17-97,
//                                         ConstrainedTimeResidual.cpp:206-413
//   state::prepareStepSystem              the assembly half of TimeIntegrator::solveStep
//                                         (femx/state/TimeIntegrator.hpp:272-289)
//
// Every operation is a C-ABI call in femx's order (CSR matvec/matvecT,
// gather/scatter, axpby, the boundary operations and the GLS kernels), so
// results equal femx's host code bit for bit.  Not built: the parameterised
// initial state (InitialStateMap), which is off the assembly path.
#pragma once

#include <algorithm>
#include <cmath>
#include <memory>
#include <set>
#include <utility>
#include <vector>

#include "femx_hip.hpp"

namespace femx
{
namespace state
{

/** Dimensions of a time-dependent residual problem (TimeResidual.hpp:13-21). */
struct TimeDims
{
  Index num_steps  = 0;
  Index num_states = 0;
  Index num_param  = 0;
  Index num_res    = 0;
  Index num_hist   = 1;
};

/** Non-owning lag-major history window (TimeResidual.hpp:23-78). */
class TimeHistoryView
{
public:
  TimeHistoryView() = default;
  TimeHistoryView(const Real* data, Index count, Index num_states)
      : data_(data), count_(count), num_states_(num_states)
  {
  }
  bool                  empty() const { return count_ == 0; }
  Index                 count() const { return count_; }
  Index                 stateSize() const { return num_states_; }
  const Real*           data() const { return data_; }
  DeviceConstVectorView state(Index i) const { return {data_ + i * num_states_, num_states_}; }
  DeviceConstVectorView values() const { return {data_, count_ * num_states_}; }

private:
  const Real* data_       = nullptr;
  Index       count_      = 0;
  Index       num_states_ = 0;
};

/** Inputs of one residual step (TimeResidual.hpp:80-86). */
struct TimeContext
{
  Index                 step{0};
  DeviceConstVectorView nxt;
  DeviceConstVectorView prm;
  TimeHistoryView       hist;
};

/** Time-residual contract over the HIP backend (TimeResidual.hpp:170-239). */
class TimeResidual
{
public:
  using Vec       = DeviceVector;
  using VecView   = DeviceVectorView;
  using ConstView = DeviceConstVectorView;
  using Mat       = DeviceCsrMatrix;
  using Pattern   = DeviceCsrPattern;
  using Ctx       = HipContext;
  using StepCtx   = TimeContext;

  virtual ~TimeResidual() = default;
  virtual TimeDims dims() const = 0;
  /** The host pattern used to construct boundary metadata. */
  virtual const HostCsrPattern& hostPattern() const = 0;
  /** The matrix pattern in device storage. */
  virtual const Pattern& pattern() const = 0;
  /** The parameter-dependent initial state. */
  virtual void initialState(ConstView prm, Vec& out, Ctx& ctx) const = 0;
  /** Adds the initial-state Jacobian transpose product to out. */
  virtual void addInitialStateJacobianTranspose(ConstView state_grad, VecView out, Ctx&) const
  {
    const TimeDims dim = dims();
    require(state_grad.size() == dim.num_states && out.size() == dim.num_param,
            "TimeResidual initial-state transpose size mismatch");
  }
  virtual void res(const StepCtx& time, Vec& out, Ctx& ctx) const = 0;
  /** The residual and the next-state Jacobian together. */
  virtual void assembleNext(const StepCtx& time, Vec& res, Mat& jac, Ctx& ctx) const = 0;
  /** A history or parameter Jacobian transpose applied to adj. */
  virtual void applyJacT(const StepCtx& time, VariableBlock wrt, ConstView adj, Vec& out,
                         Ctx& ctx) const = 0;
  virtual void prepareLinearSolve(const StepCtx&, Mat&, Vec&, Ctx&) const {}
};

} // namespace state

namespace model::ns
{

/** The residual of femx's Navier-Stokes model (NavierStokesModel::Residual,
 * femx/model/ns/NavierStokesModel.cpp:446-545) on the device: the GLS
 * NavierOperator over two history lags, parameter-free.  Its history
 * transpose is applyNavierHistJacT (femx needs Enzyme for it).  The geometry
 * and map must outlive it. */
class NavierResidual final : public state::TimeResidual
{
public:
  static constexpr Index kNumHist = 2;

  NavierResidual(const assembly::DeviceNavierGeometry& geom, const assembly::DeviceAssemblyMap& map,
                 Index num_steps, Real dt, FluidParams fluid)
      : geom_(geom), map_(map), nstep_(num_steps), op_{&geom, fluid, dt}
  {
    require(num_steps > 0, "NavierStokesModel requires at least one time step");
    require(std::isfinite(dt) && dt > 0.0, "NavierStokesModel requires a positive finite time step");
    require(std::isfinite(fluid.rho) && fluid.rho > 0.0,
            "NavierStokesModel requires positive finite density");
    require(std::isfinite(fluid.mu) && fluid.mu > 0.0,
            "NavierStokesModel requires positive finite viscosity");
    require(geom.num_elems == map.info.num_elems && geom.num_states == map.info.num_states,
            "Geometry and AssemblyMap have different element counts");
  }

  state::TimeDims dims() const override
  {
    return {nstep_, map_.info.num_states, 0, map_.info.num_res, kNumHist};
  }
  const HostCsrPattern& hostPattern() const override { return map_.pattern->host(); }
  const Pattern&        pattern() const override { return *map_.pattern; }
  const NavierOperator& op() const noexcept { return op_; }

  void initialState(ConstView prm, Vec& out, Ctx& ctx) const override
  {
    require(prm.empty(), "Navier physics residual is parameter-free");
    linalg::HipVectorHandler(ctx).resizeOrZero(out, map_.info.num_states);
  }

  void res(const StepCtx& time, Vec& out, Ctx& ctx) const override
  {
    checkCtx(time);
    assembly::assemble(op_, time.step, kNumHist, state::VariableBlock::NextState, map_,
                       history(time), time.nxt, out, nullptr, ctx);
  }

  void assembleNext(const StepCtx& time, Vec& res, Mat& jac, Ctx& ctx) const override
  {
    checkCtx(time);
    assembly::assemble(op_, time.step, kNumHist, state::VariableBlock::NextState, map_,
                       history(time), time.nxt, res, &jac, ctx);
  }

  void applyJacT(const StepCtx& time, state::VariableBlock wrt, ConstView adj, Vec& out,
                 Ctx& ctx) const override
  {
    checkCtx(time);
    require(!wrt.isNextState(), "Navier transpose apply supports only history and parameter blocks");
    require(adj.size() == map_.info.num_res, "Navier residual adjoint size mismatch");
    if (wrt.isParam())
    {
      out.resize(0);
      return;
    }
    require(wrt.historyLag() >= 0 && wrt.historyLag() < kNumHist,
            "Navier residual history lag is out of range");
    applyNavierHistJacT(op_, time.step, kNumHist, wrt.historyLag(), history(time), time.nxt, adj,
                        out, ctx);
  }

private:
  void checkCtx(const StepCtx& t) const
  {
    require(t.step >= 0 && t.step < nstep_, "Navier residual step is out of range");
    require(t.hist.count() >= kNumHist && t.hist.stateSize() == map_.info.num_states
                && t.nxt.size() == map_.info.num_states && t.prm.empty(),
            "Navier residual vector size mismatch");
  }
  DeviceConstVectorView history(const StepCtx& t) const
  {
    return {t.hist.data(), kNumHist * map_.info.num_states};
  }

  const assembly::DeviceNavierGeometry& geom_;
  const assembly::DeviceAssemblyMap&    map_;
  Index                                 nstep_;
  NavierOperator                        op_;
};

} // namespace model::ns

namespace fem
{

/** One entry of a Dirichlet control map: state row, control column, weight
 * (femx/fem/DirichletControl.hpp:14-20). */
struct DirichletControlMapEntry
{
  Index row    = 0;
  Index col    = 0;
  Real  weight = 0.0;
};

/** u_D = P q on selected state DOFs (DirichletControl.cpp:21-73): P as CSR
 * with each row's entries in input order, as makeControlMatrix builds it. */
class DirichletControl
{
public:
  DirichletControl() = default;
  /** The identity control of the given DOFs. */
  explicit DirichletControl(std::vector<Index> dofs)
  {
    std::vector<DirichletControlMapEntry> e;
    for (Index i = 0; i < static_cast<Index>(dofs.size()); ++i) e.push_back({i, i, 1.0});
    *this = DirichletControl(std::move(dofs), static_cast<Index>(e.size()), std::move(e));
  }
  DirichletControl(std::vector<Index> state_dofs, Index num_ctr_params,
                   std::vector<DirichletControlMapEntry> entries)
      : dofs_(std::move(state_dofs)), num_ctr_(num_ctr_params)
  {
    require(num_ctr_params >= 0, "DirichletControl received negative control parameter count");
    const Index                     rows = static_cast<Index>(dofs_.size());
    std::set<std::pair<Index, Index>> seen;
    std::vector<Index>              count(rows + 1, 0);
    for (const auto& e : entries)
    {
      require(e.row >= 0 && e.row < rows, "DirichletControl map state row is out of range");
      require(e.col >= 0 && e.col < num_ctr_params,
              "DirichletControl map control column is out of range");
      require(std::isfinite(e.weight), "DirichletControl received non-finite map weight");
      require(seen.insert({e.row, e.col}).second, "DirichletControl received duplicate map entry");
      ++count[e.row + 1];
    }
    for (Index r = 0; r < rows; ++r) count[r + 1] += count[r];
    std::vector<Index> next(count.begin(), count.end() - 1);
    P_.rows = rows, P_.cols = num_ctr_params, P_.row_ptr = count;
    P_.col_ind.resize(entries.size());
    vals_.resize(entries.size());
    for (const auto& e : entries)
    {
      const Index k  = next[e.row]++;
      P_.col_ind[k] = e.col;
      vals_[k]       = e.weight;
    }
    std::set<Index> uniq;
    for (Index d : dofs_)
    {
      require(d >= 0, "DirichletControl received negative state id");
      require(uniq.insert(d).second, "DirichletControl received duplicate state id");
    }
  }
  Index                     numStateDofs() const { return static_cast<Index>(dofs_.size()); }
  Index                     numControlParams() const { return num_ctr_; }
  const std::vector<Index>& stateDofs() const { return dofs_; }
  const HostCsrPattern&     matrixPattern() const { return P_; }
  const std::vector<Real>&  matrixValues() const { return vals_; }

private:
  std::vector<Index> dofs_;
  Index              num_ctr_ = 0;
  HostCsrPattern     P_;
  std::vector<Real>  vals_;
};

/** Linear time interpolation of control levels
 * (femx/common/LinearInterpolation.hpp:12-40). */
struct LinearInterpolation
{
  Index lower        = 0;
  Index upper        = 0;
  Real  upper_weight = 0.0;
  Real  lowerWeight() const { return 1.0 - upper_weight; }
  Real  upperWeight() const { return upper_weight; }
  bool  hasUpper() const { return upper != lower && upper_weight != 0.0; }
  bool  isValid() const
  {
    return lower >= 0 && upper >= lower && upper_weight >= 0.0 && upper_weight <= 1.0
           && std::isfinite(upper_weight) && (upper != lower || upper_weight == 0.0);
  }
};

/** Fixed and controlled Dirichlet data of a time-dependent problem, host side
 * (makeControlMap, ControlMap.cpp:36-137): the controlled DOFs then the fixed
 * ones, per-step fixed values, one interpolation stencil per step. */
struct HostControlMap
{
  Index                            num_steps = 0, num_states = 0, num_prm = 0, ctr_off = 0;
  std::vector<Index>               dofs; // controlled, then fixed
  Index                            num_controlled = 0, num_fixed = 0, block = 0;
  std::vector<Real>                fixed_vals; // num_steps x num_fixed
  std::vector<LinearInterpolation> time;
  HostCsrPattern                   P;
  std::vector<Real>                P_vals;
  Index                            numBcs() const { return static_cast<Index>(dofs.size()); }
};

inline HostControlMap makeControlMap(Index num_steps, Index num_states, const DirichletControl& ctr,
                                     std::vector<Index> fixed_dofs, std::vector<Real> fixed_vals,
                                     std::vector<LinearInterpolation> time, Index ctr_off,
                                     Index num_prm = -1)
{
  require(num_steps > 0 && num_states > 0 && ctr_off >= 0, "ControlMap received invalid dimensions");
  std::vector<char> used(num_states, 0);
  for (Index d : ctr.stateDofs())
  {
    require(d >= 0 && d < num_states && !used[d], "ControlMap controlled DOF is invalid");
    used[d] = 1;
  }
  for (Index d : fixed_dofs)
  {
    require(d >= 0 && d < num_states && !used[d], "ControlMap fixed DOF is invalid");
    used[d] = 1;
  }
  if (time.empty())
    for (Index s = 0; s < num_steps; ++s) time.push_back({s, s, 0.0});
  require(static_cast<Index>(time.size()) == num_steps, "ControlMap time stencil count mismatch");
  Index levels = 0;
  for (const auto& t : time)
  {
    require(t.isValid(), "ControlMap time stencil is invalid");
    levels = std::max(levels, t.upper + 1);
  }
  const Index required = ctr_off + levels * ctr.numControlParams();
  if (num_prm < 0) num_prm = required;
  require(num_prm >= required, "ControlMap parameter count is too small");
  const Index nf = static_cast<Index>(fixed_dofs.size());
  HostControlMap m;
  if (fixed_vals.empty())
    m.fixed_vals.assign(static_cast<std::size_t>(num_steps) * nf, 0.0);
  else if (static_cast<Index>(fixed_vals.size()) == nf)
    for (Index s = 0; s < num_steps; ++s)
      m.fixed_vals.insert(m.fixed_vals.end(), fixed_vals.begin(), fixed_vals.end());
  else
  {
    require(static_cast<Index>(fixed_vals.size()) == num_steps * nf,
            "ControlMap fixed value size mismatch");
    m.fixed_vals = std::move(fixed_vals);
  }
  m.num_steps = num_steps, m.num_states = num_states, m.num_prm = num_prm, m.ctr_off = ctr_off;
  m.dofs = ctr.stateDofs();
  m.dofs.insert(m.dofs.end(), fixed_dofs.begin(), fixed_dofs.end());
  m.num_controlled = ctr.numStateDofs(), m.num_fixed = nf, m.block = ctr.numControlParams();
  m.time = std::move(time);
  m.P = ctr.matrixPattern(), m.P_vals = ctr.matrixValues();
  return m;
}

/** The control map in device storage (ControlMap<Device>). */
struct DeviceControlMap
{
  HostControlMap                         host;
  std::shared_ptr<const DeviceCsrPattern> P_pattern;
  mutable DeviceCsrMatrix                P;
  DeviceIndexVector                      dofs;
  DeviceVector                           fixed_vals;
  mutable DeviceVector                   compact; // num_controlled scratch
};

/** Host -> device copy (ControlMap.hpp:146-148). */
inline void copy(const HostControlMap& src, DeviceControlMap& dst, HipContext& ctx)
{
  dst.host      = src;
  dst.P_pattern = std::make_shared<DeviceCsrPattern>(src.P, ctx);
  dst.P         = DeviceCsrMatrix(dst.P_pattern);
  dst.P.vals().assign(src.P_vals, ctx);
  dst.dofs.assign(src.dofs, ctx);
  dst.fixed_vals.assign(src.fixed_vals, ctx);
  dst.compact.resize(src.num_controlled);
}

namespace detail
{
inline const LinearInterpolation& stencil(const DeviceControlMap& m, Index step)
{
  require(step >= 0 && step < m.host.num_steps, "ControlMap vector size mismatch");
  return m.host.time[step];
}
inline DeviceConstVectorView level(const DeviceControlMap& m, DeviceConstVectorView prm, Index l)
{
  return prm.block(m.host.ctr_off + l * m.host.block, m.host.block);
}
inline DeviceVectorView level(const DeviceControlMap& m, DeviceVectorView prm, Index l)
{
  return prm.block(m.host.ctr_off + l * m.host.block, m.host.block);
}
} // namespace detail

/** Prescribed values of every constrained DOF at one step (controlVals,
 * ControlMap.cpp:165-199): the interpolated P q, then the fixed values. */
inline void controlVals(const DeviceControlMap& m, Index step, DeviceConstVectorView prm,
                        DeviceVectorView out, HipContext& ctx)
{
  require(prm.size() == m.host.num_prm && out.size() == m.host.numBcs(),
          "ControlMap vector size mismatch");
  const auto& t  = detail::stencil(m, step);
  auto        ct = out.block(0, m.host.num_controlled);
  linalg::HipMatrixHandler mh(ctx);
  mh.matvec(m.P, detail::level(m, prm, t.lower), ct, t.lowerWeight(), 0.0);
  if (t.hasUpper()) mh.matvec(m.P, detail::level(m, prm, t.upper), ct, t.upperWeight(), 1.0);
  if (m.host.num_fixed > 0)
    linalg::HipVectorHandler(ctx).copy(
        DeviceConstVectorView(m.fixed_vals).block(step * m.host.num_fixed, m.host.num_fixed),
        out.block(m.host.num_controlled, m.host.num_fixed));
}

/** d(residual)/d(parameters) applied to a direction (controlJac,
 * ControlMap.cpp:237-271): -P (interpolated direction) on the controlled
 * DOFs, zero elsewhere. */
inline void controlJac(const DeviceControlMap& m, Index step, DeviceConstVectorView dir,
                       DeviceVectorView out, HipContext& ctx)
{
  require(dir.size() == m.host.num_prm && out.size() == m.host.num_states,
          "ControlMap vector size mismatch");
  const auto&              t = detail::stencil(m, step);
  linalg::HipVectorHandler vh(ctx);
  linalg::HipMatrixHandler mh(ctx);
  vh.zero(out);
  mh.matvec(m.P, detail::level(m, dir, t.lower), view(m.compact), -t.lowerWeight(), 0.0);
  if (t.hasUpper()) mh.matvec(m.P, detail::level(m, dir, t.upper), view(m.compact), -t.upperWeight(), 1.0);
  vh.scatter(view(m.compact), DeviceConstIndexView(m.dofs).block(0, m.host.num_controlled), out);
}

/** grad += (d residual / d parameters)^T adj (addControlJacT,
 * ControlMap.cpp:309-342). */
inline void addControlJacT(const DeviceControlMap& m, Index step, DeviceConstVectorView adj,
                           DeviceVectorView grad, HipContext& ctx)
{
  require(adj.size() == m.host.num_states && grad.size() == m.host.num_prm,
          "ControlMap vector size mismatch");
  const auto&              t = detail::stencil(m, step);
  linalg::HipVectorHandler vh(ctx);
  linalg::HipMatrixHandler mh(ctx);
  vh.gather(adj, DeviceConstIndexView(m.dofs).block(0, m.host.num_controlled), view(m.compact));
  mh.matvecT(m.P, view(m.compact), detail::level(m, grad, t.lower), -t.lowerWeight(), 1.0);
  if (t.hasUpper())
    mh.matvecT(m.P, view(m.compact), detail::level(m, grad, t.upper), -t.upperWeight(), 1.0);
}

} // namespace fem

namespace assembly
{

/** A time residual decorated with Dirichlet constraints, device storage
 * (ConstrainedTimeResidual<CudaCsrBackend>, femx/assembly/
 * ConstrainedTimeResidual.cpp:232-405): constrained rows of the residual
 * become nxt - u_D, of the Jacobian identity rows, and the adjoint is zeroed
 * there before a history transpose.  It owns the base residual, as femx's
 * device form does.  Calls must not overlap (shared scratch vectors). */
class ConstrainedTimeResidual final : public state::TimeResidual
{
public:
  ConstrainedTimeResidual(std::unique_ptr<state::TimeResidual> base, fem::HostControlMap control,
                          HipContext& ctx)
      : base_(std::move(base))
  {
    require(base_ != nullptr, "ConstrainedTimeResidual requires a base residual");
    base_dims_ = base_->dims();
    require(base_dims_.num_res == base_dims_.num_states,
            "ConstrainedTimeResidual requires square state residuals");
    require(control.num_steps == base_dims_.num_steps
                && control.num_states == base_dims_.num_states,
            "ConstrainedTimeResidual control dimensions do not match");
    dims_           = base_dims_;
    dims_.num_param = base_dims_.num_param + control.num_prm - base_dims_.num_param;
    require(base_dims_.num_param == 0,
            "ConstrainedTimeResidual requires a parameter-free base residual");
    const auto hb = makeBoundaryMap(control.dofs, base_->hostPattern());
    copy(hb, boundary_, ctx);
    fem::copy(control, control_, ctx);
    base_adj_.resize(dims_.num_res);
    boundary_vals_.resize(control.numBcs());
  }

  state::TimeDims          dims() const override { return dims_; }
  const HostCsrPattern&    hostPattern() const override { return base_->hostPattern(); }
  const Pattern&           pattern() const override { return base_->pattern(); }
  const fem::DeviceControlMap& controlMap() const noexcept { return control_; }

  void initialState(ConstView prm, Vec& out, Ctx& ctx) const override
  {
    require(prm.size() == dims_.num_param,
            "ConstrainedTimeResidual initial-state parameter size mismatch");
    base_->initialState(ConstView{}, out, ctx);
  }

  void res(const StepCtx& time, Vec& out, Ctx& ctx) const override
  {
    checkCtx(time);
    base_->res(baseCtx(time), out, ctx);
    require(out.size() == dims_.num_res, "ConstrainedTimeResidual base residual size mismatch");
    fem::controlVals(control_, time.step, time.prm, view(boundary_vals_), ctx);
    replaceRes(boundary_, time.nxt, view(boundary_vals_), view(out), ctx);
  }

  void assembleNext(const StepCtx& time, Vec& res, Mat& jac, Ctx& ctx) const override
  {
    checkCtx(time);
    base_->assembleNext(baseCtx(time), res, jac, ctx);
    require(res.size() == dims_.num_res, "ConstrainedTimeResidual base residual size mismatch");
    fem::controlVals(control_, time.step, time.prm, view(boundary_vals_), ctx);
    replaceRes(boundary_, time.nxt, view(boundary_vals_), view(res), ctx);
    replaceRows(boundary_, jac, 1.0, ctx);
  }

  void applyJacT(const StepCtx& time, state::VariableBlock wrt, ConstView adj, Vec& out,
                 Ctx& ctx) const override
  {
    checkCtx(time);
    require(!wrt.isNextState(),
            "Constrained transpose apply supports only history and parameter blocks");
    require(adj.size() == dims_.num_res, "ConstrainedTimeResidual adjoint size mismatch");
    linalg::HipVectorHandler vh(ctx);
    if (wrt.isParam())
    {
      vh.resizeOrZero(out, dims_.num_param);
      fem::addControlJacT(control_, time.step, adj, view(out), ctx);
      return;
    }
    vh.copy(adj, view(base_adj_));
    zeroBoundary(boundary_, base_adj_, ctx);
    base_->applyJacT(baseCtx(time), wrt, view(base_adj_), out, ctx);
    require(out.size() == dims_.num_states, "ConstrainedTimeResidual transpose apply size mismatch");
  }

  void prepareLinearSolve(const StepCtx& time, Mat& jac, Vec& rhs, Ctx& ctx) const override
  {
    checkCtx(time);
    base_->prepareLinearSolve(baseCtx(time), jac, rhs, ctx);
    fem::controlVals(control_, time.step, time.prm, view(boundary_vals_), ctx);
    prepareForwardSolve(boundary_, jac, rhs, boundary_vals_, ctx);
  }

private:
  StepCtx baseCtx(const StepCtx& time) const
  {
    StepCtx b = time;
    b.prm     = ConstView{}; // the base residual sees its own (empty) parameters
    return b;
  }
  void checkCtx(const StepCtx& t) const
  {
    require(t.step >= 0 && t.step < dims_.num_steps && t.hist.count() >= dims_.num_hist
                && t.hist.stateSize() == dims_.num_states && t.nxt.size() == dims_.num_states
                && t.prm.size() == dims_.num_param,
            "ConstrainedTimeResidual context dimensions do not match");
  }

  std::unique_ptr<state::TimeResidual> base_;
  fem::DeviceControlMap                control_;
  DeviceBoundaryMap                    boundary_;
  mutable DeviceVector                 base_adj_, boundary_vals_;
  state::TimeDims                      base_dims_, dims_;
};

} // namespace assembly

namespace model::ns
{
/** makeDeviceTimeResidual (femx/model/ns/NavierStokesModel.hpp:126-129): the
 * model residual under ConstrainedTimeResidual. */
inline std::unique_ptr<state::TimeResidual>
makeDeviceTimeResidual(std::unique_ptr<NavierResidual> model, fem::HostControlMap control,
                       HipContext& ctx)
{
  return std::make_unique<assembly::ConstrainedTimeResidual>(std::move(model), std::move(control),
                                                             ctx);
}
} // namespace model::ns

namespace state
{
/** The assembly half of TimeIntegrator::solveStep (femx/state/
 * TimeIntegrator.hpp:272-289): assembleNext, rhs = J nxt - res, then
 * prepareLinearSolve.  The linear solve is the caller's. */
inline void prepareStepSystem(const TimeResidual& r, const TimeContext& time, DeviceVector& res,
                              DeviceCsrMatrix& jac, DeviceVector& rhs, HipContext& ctx)
{
  r.assembleNext(time, res, jac, ctx);
  linalg::HipMatrixHandler(ctx).matvec(jac, time.nxt, rhs);
  linalg::HipVectorHandler(ctx).axpby(-1.0, view(res), 1.0, view(rhs));
  r.prepareLinearSolve(time, jac, rhs, ctx);
}
} // namespace state

} // namespace femx
