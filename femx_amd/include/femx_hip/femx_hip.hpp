// femx_hip.hpp -- femx-shaped C++ host interface over the femx_amd C-ABI.
//
// Header-only.  Mirrors the reference's host types and free functions for
// the FE hot path so that a femx caller swaps the CUDA overloads for these:
//
//   femx::HipContext                 femx/common/Context.hpp:48-90 (CudaContext)
//   femx::DeviceArray<T>             femx/linalg/Vector.hpp:322-400 (device Vector)
//   femx::DeviceCsrPattern/Matrix    femx/linalg/CsrPattern.hpp:43-217, CsrMatrix.hpp:17-105
//   assembly::HostAssemblyMap,       femx/assembly/AssemblyMap.hpp:75-246
//     makeAssemblyMap, copy
//   assembly::HostBoundaryMap,       femx/assembly/BoundaryMap.hpp:71-294
//     makeBoundaryMap, copy, replaceRows, replaceRes, zeroBoundary,
//     prepareForwardSolve
//   assembly::assemble               femx/assembly/CudaAssembly.hpp:319-364
//   linalg::HipMatrixHandler         femx/linalg/handler/MatrixHandler.hpp:193-378
//   linalg::HipVectorHandler         femx/linalg/handler/VectorHandler.hpp:190-477
//   femx::VectorView                 femx/linalg/View.hpp:13-90
//   state::Residual                  femx/state/Residual.hpp:16-60
//   state::PoissonLogKappaResidual   a Residual for the P1 log-kappa inverse problem
//   state::NavierStokesResidual      a Residual for the Taylor-Hood nu_e inverse problem
//
// femx_hip/time.hpp holds the time-residual contract, femx's Navier-Stokes
// model residual, Dirichlet control and ConstrainedTimeResidual;
// femx_hip/solver.hpp the LinearSolver contract and a device BiCGStab.
//
// Errors surface as std::runtime_error carrying the library's message, as
// femx's require() does (femx/common/Checks.hpp:11-17).  Device work is
// enqueued on the context's stream and is asynchronous, like femx's CUDA
// path; call HipContext::sync() before reading results on the host.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/femx_amd.h"

namespace femx
{

using Real  = double;
using Index = std::int32_t;

inline void check(int status)
{
  if (status != FX_OK) throw std::runtime_error(fx_last_error());
}

inline void checkHip(hipError_t st, const char* what)
{
  if (st != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(st));
}

inline void require(bool cond, const std::string& msg)
{
  if (!cond) throw std::runtime_error(msg);
}

/** Owns one HIP stream (default priority), like femx's CudaContext. */
class HipContext
{
public:
  HipContext() { checkHip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate"); }
  ~HipContext()
  {
    if (stream_) (void)hipStreamDestroy(stream_);
  }
  HipContext(const HipContext&)            = delete;
  HipContext& operator=(const HipContext&) = delete;

  static bool available()
  {
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
  }
  hipStream_t stream() const noexcept { return stream_; }
  void*       raw() const noexcept { return stream_; }
  void        sync() const { checkHip(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

private:
  hipStream_t stream_ = nullptr;
};

/** Move-only device array (femx's Vector<Device, T>). */
template <class T>
class DeviceArray
{
public:
  DeviceArray() = default;
  explicit DeviceArray(Index n) { resize(n); }
  DeviceArray(const std::vector<T>& host, HipContext& ctx) { assign(host, ctx); }
  ~DeviceArray() { release(); }
  DeviceArray(DeviceArray&& o) noexcept : data_(std::exchange(o.data_, nullptr)), size_(std::exchange(o.size_, 0)) {}
  DeviceArray& operator=(DeviceArray&& o) noexcept
  {
    if (this != &o)
    {
      release();
      data_ = std::exchange(o.data_, nullptr);
      size_ = std::exchange(o.size_, 0);
    }
    return *this;
  }
  DeviceArray(const DeviceArray&)            = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;

  void resize(Index n)
  {
    require(n >= 0, "DeviceArray size must be non-negative");
    if (n == size_) return;
    release();
    if (n > 0) checkHip(hipMalloc(&data_, sizeof(T) * static_cast<std::size_t>(n)), "hipMalloc");
    size_ = n;
  }
  void assign(const std::vector<T>& host, HipContext& ctx)
  {
    resize(static_cast<Index>(host.size()));
    if (size_ > 0)
      checkHip(hipMemcpyAsync(data_, host.data(), sizeof(T) * host.size(), hipMemcpyHostToDevice,
                              ctx.stream()),
               "hipMemcpyAsync");
  }
  std::vector<T> toHost(HipContext& ctx) const
  {
    std::vector<T> out(size_);
    if (size_ > 0)
      checkHip(hipMemcpyAsync(out.data(), data_, sizeof(T) * out.size(), hipMemcpyDeviceToHost,
                              ctx.stream()),
               "hipMemcpyAsync");
    ctx.sync();
    return out;
  }
  T*       data() noexcept { return data_; }
  const T* data() const noexcept { return data_; }
  Index    size() const noexcept { return size_; }

private:
  void release()
  {
    if (data_) (void)hipFree(data_);
    data_ = nullptr;
    size_ = 0;
  }
  T*    data_ = nullptr;
  Index size_ = 0;
};

using DeviceVector      = DeviceArray<Real>;
using DeviceIndexVector = DeviceArray<Index>;

/** Non-owning view of contiguous device values (femx's VectorView<Device, T>,
 * femx/linalg/View.hpp:13-90): a pointer and a size. */
template <class T>
class VectorView
{
public:
  VectorView() = default;
  VectorView(T* data, Index size) : data_(data), size_(size) {}
  template <class U>
  VectorView(const VectorView<U>& o) : data_(o.data()), size_(o.size())
  {
  }
  template <class U>
  VectorView(DeviceArray<U>& a) : data_(a.data()), size_(a.size())
  {
  }
  template <class U>
  VectorView(const DeviceArray<U>& a) : data_(a.data()), size_(a.size())
  {
  }
  T*    data() const noexcept { return data_; }
  Index size() const noexcept { return size_; }
  bool  empty() const noexcept { return size_ == 0; }
  bool  isValid() const noexcept { return size_ >= 0 && (size_ == 0 || data_ != nullptr); }
  /** The sub-range [off, off + n) (femx's BlockVectorView::block). */
  VectorView block(Index off, Index n) const
  {
    require(off >= 0 && n >= 0 && off + n <= size_, "VectorView block out of range");
    return {data_ + off, n};
  }

private:
  T*    data_ = nullptr;
  Index size_ = 0;
};

using DeviceVectorView      = VectorView<Real>;
using DeviceConstVectorView = VectorView<const Real>;
using DeviceIndexView       = VectorView<Index>;
using DeviceConstIndexView  = VectorView<const Index>;

/** Views of a device array (femx's Vector::view). */
template <class T>
VectorView<T> view(DeviceArray<T>& a)
{
  return {a.data(), a.size()};
}
template <class T>
VectorView<const T> view(const DeviceArray<T>& a)
{
  return {a.data(), a.size()};
}

namespace detail
{
inline void fill(Real* y, Index n, Real v, void* stream) { check(fx_vec_fill(n, v, y, stream)); }

inline bool overlap(const void* a, std::size_t na, const void* b, std::size_t nb)
{
  if (!a || !b || na == 0 || nb == 0) return false;
  auto pa = reinterpret_cast<std::uintptr_t>(a), pb = reinterpret_cast<std::uintptr_t>(b);
  return pa < pb + nb && pb < pa + na;
}
} // namespace detail

namespace detail
{
inline std::uint64_t newLayoutId()
{
  static std::atomic<std::uint64_t> next{1};
  return next++;
}
} // namespace detail

/** Immutable host CSR pattern with a layout identity. */
struct HostCsrPattern
{
  Index              rows = 0, cols = 0;
  std::vector<Index> row_ptr{0}, col_ind;
  std::uint64_t      layout_id = detail::newLayoutId();
  Index              nnz() const { return static_cast<Index>(col_ind.size()); }
};

/** Device CSR pattern sharing the host pattern's layout identity. */
class DeviceCsrPattern
{
public:
  DeviceCsrPattern() = default;
  DeviceCsrPattern(const HostCsrPattern& h, HipContext& ctx)
    : rows_(h.rows), cols_(h.cols), layout_id_(h.layout_id), host_(h)
  {
    row_ptr_.assign(h.row_ptr, ctx);
    col_ind_.assign(h.col_ind, ctx);
  }
  Index                 rows() const noexcept { return rows_; }
  Index                 cols() const noexcept { return cols_; }
  Index                 nnz() const noexcept { return col_ind_.size(); }
  std::uint64_t         layoutId() const noexcept { return layout_id_; }
  const Index*          rowPtrData() const noexcept { return row_ptr_.data(); }
  const Index*          colIndData() const noexcept { return col_ind_.data(); }
  const HostCsrPattern& host() const noexcept { return host_; }

  struct TransposeMap
  {
    DeviceIndexVector t_ptr, t_row, t_perm;
  };

  /** Cached transpose permutation (t_ptr, t_row, t_perm) on the device. */
  const TransposeMap& transposeMap(HipContext& ctx) const
  {
    if (!tmap_)
    {
      std::vector<Index> tp(cols_ + 1), tr(nnz()), tq(nnz());
      check(fx_csr_transpose_map(rows_, cols_, host_.row_ptr.data(), host_.col_ind.data(), tp.data(),
                                 tr.data(), tq.data()));
      auto m = std::make_shared<TransposeMap>();
      m->t_ptr.assign(tp, ctx);
      m->t_row.assign(tr, ctx);
      m->t_perm.assign(tq, ctx);
      tmap_ = std::move(m);
      HostCsrPattern t; // columns become rows, each row's columns ascending
      t.rows = cols_, t.cols = rows_, t.row_ptr = std::move(tp), t.col_ind = std::move(tr);
      tpat_ = std::make_shared<DeviceCsrPattern>(t, ctx);
    }
    return *tmap_;
  }

  /** The transposed pattern (a layout of its own), built with the
   * transpose map and cached with it (MatrixHandler.cu:401-512). */
  std::shared_ptr<const DeviceCsrPattern> transposed(HipContext& ctx) const
  {
    transposeMap(ctx);
    return tpat_;
  }

  /** Cached 16-bit column offsets (col_ind - row) on the device, or null when
   * the pattern's bandwidth does not fit 16 bits (fx_csr_col_off16). */
  const DeviceArray<std::int16_t>* colOff16(HipContext& ctx) const
  {
    if (!off16_tried_)
    {
      off16_tried_ = true;
      std::vector<std::int16_t> off(nnz());
      if (fx_csr_col_off16(rows_, host_.row_ptr.data(), host_.col_ind.data(), off.data()) == FX_OK)
      {
        auto d = std::make_shared<DeviceArray<std::int16_t>>();
        d->assign(off, ctx);
        off16_ = std::move(d);
      }
    }
    return off16_.get();
  }

private:
  Index                                             rows_ = 0, cols_ = 0;
  std::uint64_t                                     layout_id_ = 0;
  HostCsrPattern                                    host_;
  DeviceIndexVector                                 row_ptr_, col_ind_;
  mutable std::shared_ptr<TransposeMap>             tmap_;
  mutable std::shared_ptr<const DeviceCsrPattern>   tpat_;
  mutable std::shared_ptr<DeviceArray<std::int16_t>> off16_;
  mutable bool                                      off16_tried_ = false;
};

/** Device CSR matrix: shared pattern, owned values. */
class DeviceCsrMatrix
{
public:
  DeviceCsrMatrix() = default;
  explicit DeviceCsrMatrix(std::shared_ptr<const DeviceCsrPattern> p) : pattern_(std::move(p))
  {
    vals_.resize(pattern_->nnz());
  }
  const DeviceCsrPattern& pattern() const { return *pattern_; }
  bool                    hasPattern() const noexcept { return pattern_ != nullptr; }
  std::shared_ptr<const DeviceCsrPattern> sharedPattern() const { return pattern_; }
  Index        rows() const { return pattern_->rows(); }
  Index        cols() const { return pattern_->cols(); }
  Index        nnz() const { return pattern_->nnz(); }
  Real*        valsData() { return vals_.data(); }
  const Real*  valsData() const { return vals_.data(); }
  DeviceVector&       vals() { return vals_; }
  const DeviceVector& vals() const { return vals_; }

private:
  std::shared_ptr<const DeviceCsrPattern> pattern_;
  DeviceVector                            vals_;
};

namespace assembly
{

/** Host assembly map built by fx_map_create (femx makeAssemblyMap). */
class HostAssemblyMap
{
public:
  HostAssemblyMap() = default;
  HostAssemblyMap(Index num_res, Index num_states, const std::vector<std::vector<Index>>& res_dofs,
                  const std::vector<std::vector<Index>>& state_dofs)
  {
    require(res_dofs.size() == state_dofs.size(), "AssemblyMap residual/state element counts differ");
    std::vector<Index> ro{0}, rd, so{0}, sd;
    for (std::size_t e = 0; e < res_dofs.size(); ++e)
    {
      rd.insert(rd.end(), res_dofs[e].begin(), res_dofs[e].end());
      sd.insert(sd.end(), state_dofs[e].begin(), state_dofs[e].end());
      ro.push_back(static_cast<Index>(rd.size()));
      so.push_back(static_cast<Index>(sd.size()));
    }
    fx_map* raw = nullptr;
    check(fx_map_create(num_res, num_states, static_cast<Index>(res_dofs.size()), ro.data(),
                        rd.data(), so.data(), sd.data(), &raw));
    h_.reset(raw, fx_map_destroy);
    check(fx_map_info_get(raw, &info_));
    auto p     = std::make_shared<HostCsrPattern>();
    p->rows    = info_.num_res;
    p->cols    = info_.num_states;
    p->row_ptr.resize(info_.num_res + 1);
    p->col_ind.resize(info_.nnz);
    check(fx_map_export_pattern(raw, p->row_ptr.data(), p->col_ind.data()));
    pattern_ = std::move(p);
  }
  Index                 numElems() const { return info_.num_elems; }
  Index                 numRes() const { return info_.num_res; }
  Index                 numStates() const { return info_.num_states; }
  Index                 maxRes() const { return info_.max_res; }
  Index                 maxState() const { return info_.max_state; }
  Index                 maxJac() const { return info_.max_jac; }
  const fx_map_info&    info() const { return info_; }
  const HostCsrPattern& pattern() const { return *pattern_; }
  const fx_map*         handle() const { return h_.get(); }

  /** femx's flattened jac_offsets / jac_map. */
  void jacMap(std::vector<Index>& offsets, std::vector<Index>& map) const
  {
    offsets.resize(info_.num_elems + 1);
    map.resize(info_.jac_len);
    check(fx_map_export_jac(h_.get(), offsets.data(), map.data()));
  }

private:
  std::shared_ptr<fx_map>               h_;
  fx_map_info                           info_{};
  std::shared_ptr<const HostCsrPattern> pattern_;
};

inline HostAssemblyMap makeAssemblyMap(Index num_res, Index num_states,
                                       const std::vector<std::vector<Index>>& res_dofs,
                                       const std::vector<std::vector<Index>>& state_dofs)
{
  return HostAssemblyMap(num_res, num_states, res_dofs, state_dofs);
}

/** Device map: pattern plus the row-gather tables of the element kernels. */
struct DeviceAssemblyMap
{
  fx_map_info                             info{};
  std::shared_ptr<const DeviceCsrPattern> pattern;
  DeviceIndexVector inc_off, inc, inc_elem, inc_meta, blk_row, blk_flags, blk_lanes;
  DeviceIndexVector elem_inc; // 15x15 maps: (element, local row) -> row incidence (the VJP)
  DeviceIndexVector rt_off, rt_elem, rt_meta; // 15x15 maps: round tables (fx_map_th_rounds)
  Index             num_blocks = 0, max_block_nnz = 0;
  // 3x3 maps: the 64-row stencil blocks of the P1 assembly (fx_p1_stencil_plan)
  DeviceIndexVector p1_blk_code, p1_words;
  double            p1_stencil_frac = 0.0; // share of rows in uniform blocks
  // Taylor-Hood VJP workspace (15 doubles per element), sized on first use;
  // VJPs of one map must not overlap
  mutable DeviceVector vjp_work;
};

/** Host -> device copy (femx copy(map, map, ctx)); builds the row-incidence
 * tables the element shape needs (3x3: P1, 15x15: Taylor-Hood). */
inline void copy(const HostAssemblyMap& src, DeviceAssemblyMap& dst, HipContext& ctx)
{
  dst.info    = src.info();
  dst.pattern = std::make_shared<DeviceCsrPattern>(src.pattern(), ctx);
  std::vector<Index> off(src.numRes() + 1);
  if (src.maxRes() == 3 && src.maxState() == 3)
  {
    std::vector<Index> inc(2 * static_cast<std::size_t>(src.info().num_inc));
    check(fx_map_export_rowinc3(src.handle(), off.data(), inc.data()));
    dst.inc_off.assign(off, ctx);
    dst.inc.assign(inc, ctx);
    const HostCsrPattern& hp = src.pattern();
    Index                 nb = 0, nw = 0;
    check(fx_p1_stencil_plan(src.numRes(), hp.row_ptr.data(), hp.col_ind.data(), off.data(),
                             inc.data(), FX_P1_STENCIL_MAX, FX_P1_STENCIL_BAND, &nb, &nw, nullptr,
                             nullptr));
    std::vector<Index> code(nb), words(nw);
    check(fx_p1_stencil_plan(src.numRes(), hp.row_ptr.data(), hp.col_ind.data(), off.data(),
                             inc.data(), FX_P1_STENCIL_MAX, FX_P1_STENCIL_BAND, &nb, &nw,
                             code.data(), words.data()));
    std::size_t uniform = 0;
    for (Index c : code) uniform += c >= 0;
    dst.p1_stencil_frac = src.numRes() > 0 ? 64.0 * uniform / src.numRes() : 0.0;
    if (uniform > 0)
    {
      dst.p1_blk_code.assign(code, ctx);
      dst.p1_words.assign(words, ctx);
    }
  }
  else if (src.maxRes() == 15 && src.maxState() == 15)
  {
    std::vector<Index> el(src.info().num_inc), me(4 * static_cast<std::size_t>(src.info().num_inc));
    check(fx_map_export_rowinc15(src.handle(), off.data(), el.data(), me.data()));
    Index nb = 0, mx = 0;
    check(fx_map_th_blocks(src.handle(), FX_TH_BLOCK_ITEMS, FX_TH_TILE_NNZ, &nb, &mx, nullptr,
                           nullptr));
    std::vector<Index> br(nb + 1), bf(nb);
    check(fx_map_th_blocks(src.handle(), FX_TH_BLOCK_ITEMS, FX_TH_TILE_NNZ, &nb, &mx, br.data(),
                           bf.data()));
    std::vector<Index> bl(64 * static_cast<std::size_t>(nb));
    check(fx_map_th_lanes(src.handle(), nb, br.data(), bf.data(), bl.data()));
    std::vector<Index> ro(nb + 1);
    check(fx_map_th_rounds(src.numRes(), nb, br.data(), bf.data(), bl.data(), off.data(), el.data(), me.data(),
                           ro.data(), nullptr, nullptr));
    std::vector<Index> re(std::max<Index>(ro[nb], 1)), rm(4 * re.size());
    check(fx_map_th_rounds(src.numRes(), nb, br.data(), bf.data(), bl.data(), off.data(), el.data(), me.data(),
                           ro.data(), re.data(), rm.data()));
    dst.rt_off.assign(ro, ctx);
    dst.rt_elem.assign(re, ctx);
    dst.rt_meta.assign(rm, ctx);
    dst.inc_off.assign(off, ctx);
    dst.inc_elem.assign(el, ctx);
    dst.inc_meta.assign(me, ctx);
    dst.blk_row.assign(br, ctx);
    dst.blk_flags.assign(bf, ctx);
    dst.blk_lanes.assign(bl, ctx);
    dst.num_blocks    = nb;
    dst.max_block_nnz = mx;
    std::vector<Index> ei(15 * static_cast<std::size_t>(src.numElems()));
    check(fx_map_export_rowinc(src.handle(), off.data(), nullptr, ei.data()));
    dst.elem_inc.assign(ei, ctx);
  }
}

/** Host boundary map (femx makeBoundaryMap) with its device copy. */
struct HostBoundaryMap
{
  Index              rows = 0;
  std::uint64_t      layout_id = 0;
  std::vector<Index> bc_rows, diag, col_offsets, col_entries, col_rows, bc_mask;
  std::vector<Index> elim_rows, elim_off, elim_k, elim_ib;
  Index              numBcs() const { return static_cast<Index>(bc_rows.size()); }
};

inline HostBoundaryMap makeBoundaryMap(const std::vector<Index>& dofs, const HostCsrPattern& pattern)
{
  require(pattern.rows == pattern.cols, "BoundaryMap requires a square CSR pattern");
  fx_bcmap* raw = nullptr;
  check(fx_bcmap_create(pattern.rows, pattern.row_ptr.data(), pattern.col_ind.data(),
                        static_cast<Index>(dofs.size()), dofs.data(), &raw));
  std::unique_ptr<fx_bcmap, void (*)(fx_bcmap*)> h(raw, fx_bcmap_destroy);
  fx_bcmap_info info{};
  check(fx_bcmap_info_get(raw, &info));
  HostBoundaryMap m;
  m.rows      = pattern.rows;
  m.layout_id = pattern.layout_id;
  m.bc_rows.resize(info.num_bcs);
  m.diag.resize(info.num_bcs);
  m.col_offsets.resize(info.num_bcs + 1);
  m.col_entries.resize(info.num_col_entries);
  m.col_rows.resize(info.num_col_entries);
  m.bc_mask.resize(pattern.rows);
  check(fx_bcmap_export(raw, m.bc_rows.data(), m.diag.data(), m.col_offsets.data(),
                        m.col_entries.data(), m.col_rows.data(), m.bc_mask.data()));
  m.elim_rows.resize(info.num_elim_rows);
  m.elim_off.resize(info.num_elim_rows + 1);
  m.elim_k.resize(info.num_elim_entries);
  m.elim_ib.resize(info.num_elim_entries);
  check(fx_bcmap_export_elim(raw, m.elim_rows.data(), m.elim_off.data(), m.elim_k.data(),
                             m.elim_ib.data()));
  return m;
}

struct DeviceBoundaryMap
{
  Index             rows = 0, num_bcs = 0, num_elim_rows = 0;
  std::uint64_t     layout_id = 0;
  DeviceIndexVector bc_rows, diag, bc_mask, elim_rows, elim_off, elim_k, elim_ib;
};

inline void copy(const HostBoundaryMap& src, DeviceBoundaryMap& dst, HipContext& ctx)
{
  dst.rows          = src.rows;
  dst.num_bcs       = src.numBcs();
  dst.num_elim_rows = static_cast<Index>(src.elim_rows.size());
  dst.layout_id     = src.layout_id;
  dst.bc_rows.assign(src.bc_rows, ctx);
  dst.diag.assign(src.diag, ctx);
  dst.bc_mask.assign(src.bc_mask, ctx);
  dst.elim_rows.assign(src.elim_rows, ctx);
  dst.elim_off.assign(src.elim_off, ctx);
  dst.elim_k.assign(src.elim_k, ctx);
  dst.elim_ib.assign(src.elim_ib, ctx);
}

inline void checkLayout(const DeviceBoundaryMap& map, const DeviceCsrMatrix& mat)
{
  require(mat.pattern().layoutId() == map.layout_id,
          "BoundaryMap matrix does not match the mapped CSR layout");
}

inline void replaceRows(const DeviceBoundaryMap& map, DeviceCsrMatrix& jac, Real diag, HipContext& ctx)
{
  checkLayout(map, jac);
  check(fx_bc_replace_rows(map.num_bcs, map.bc_rows.data(), map.diag.data(),
                           jac.pattern().rowPtrData(), jac.valsData(), diag, ctx.raw()));
}

inline void replaceRes(const DeviceBoundaryMap& map, DeviceConstVectorView state,
                       DeviceConstVectorView bc_vals, DeviceVectorView res, HipContext& ctx)
{
  require(state.size() == map.rows && res.size() == map.rows && bc_vals.size() == map.num_bcs,
          "BoundaryMap residual vectors have incompatible sizes");
  check(fx_bc_replace_res(map.num_bcs, map.bc_rows.data(), state.data(), bc_vals.data(), res.data(),
                          ctx.raw()));
}

inline void zeroBoundary(const DeviceBoundaryMap& map, DeviceVectorView vals, HipContext& ctx)
{
  require(vals.size() == map.rows, "BoundaryMap vector has incompatible size");
  check(fx_bc_zero(map.num_bcs, map.bc_rows.data(), vals.data(), ctx.raw()));
}

inline void prepareForwardSolve(const DeviceBoundaryMap& map, DeviceCsrMatrix& mat,
                                DeviceVectorView rhs, DeviceConstVectorView bc_vals, HipContext& ctx)
{
  checkLayout(map, mat);
  require(rhs.size() == map.rows && bc_vals.size() == map.num_bcs,
          "BoundaryMap forward vectors have incompatible sizes");
  check(fx_bc_prepare_forward(map.num_bcs, map.bc_rows.data(), map.diag.data(),
                              mat.pattern().rowPtrData(), map.num_elim_rows, map.elim_rows.data(),
                              map.elim_off.data(), map.elim_k.data(), map.elim_ib.data(),
                              mat.valsData(), rhs.data(), bc_vals.data(), ctx.raw()));
}

/** Structure-of-arrays 2-D geometry on the device (femx Geometry). */
struct DeviceGeometry
{
  Index             num_nodes = 0, num_elems = 0, nodes_per_elem = 0;
  DeviceVector      x, y;
  DeviceIndexVector conn;
};

/** P1 Poisson operator: kappa[E] (or its log) and nodal load f[N]; null = 1 / 0. */
struct PoissonP1
{
  const Real* kappa     = nullptr;
  bool        log_kappa = false;
  const Real* f         = nullptr;
};

/** Taylor-Hood NS operator: scalar nu or per-element nu_e[E]. */
struct TaylorHoodNS
{
  Real        nu   = 0.01;
  const Real* nu_e = nullptr;
};

inline void checkAssembly(const DeviceAssemblyMap& map, const DeviceVector* state,
                          const DeviceVector& res, const DeviceCsrMatrix* jac)
{
  if (state) require(state->size() == map.info.num_states, "Assembly state size does not match AssemblyMap");
  require(res.size() == map.info.num_res, "Assembly residual size does not match AssemblyMap");
  if (jac)
  {
    require(jac->pattern().layoutId() == map.pattern->layoutId(),
            "Assembly matrix must use the AssemblyMap CSR layout");
    require(jac->valsData() != res.data() && (!state || state->data() != res.data()),
            "Assembly state, residual, and matrix values must not alias");
  }
}

/** assemble(op, geom, map, state, res, jac, ctx): residual and CSR Jacobian
 * values, both replaced (CudaAssembly.hpp:319-364).  jac may be null.  When
 * at least half of the map's rows lie in 64-row stencil blocks (built by
 * copy(map)) the stencil-block kernel runs, else the row-per-lane one; the
 * results are bitwise equal. */
inline void assemble(const PoissonP1& op, const DeviceGeometry& geom, const DeviceAssemblyMap& map,
                     const DeviceVector* state, DeviceVector& res, DeviceCsrMatrix* jac,
                     HipContext& ctx)
{
  checkAssembly(map, state, res, jac);
  require(geom.num_elems == map.info.num_elems, "Geometry and AssemblyMap have different element counts");
  if (map.p1_stencil_frac >= 0.5)
  {
    check(fx_p1_assemble_stencil(
        geom.num_nodes, geom.num_elems, map.info.nnz, map.info.max_row_len, geom.x.data(),
        geom.y.data(), geom.conn.data(), op.kappa, op.log_kappa ? 1 : 0, op.f,
        state ? state->data() : nullptr, map.inc_off.data(), map.inc.data(),
        map.pattern->rowPtrData(), map.p1_blk_code.data(), map.p1_words.data(),
        jac ? jac->valsData() : nullptr, res.data(), ctx.raw()));
    return;
  }
  check(fx_p1_assemble(geom.num_nodes, geom.num_elems, map.info.nnz, map.info.max_row_len,
                       geom.x.data(), geom.y.data(), geom.conn.data(), op.kappa, op.log_kappa ? 1 : 0,
                       op.f, state ? state->data() : nullptr, map.inc_off.data(), map.inc.data(),
                       map.pattern->rowPtrData(), jac ? jac->valsData() : nullptr, res.data(),
                       ctx.raw()));
}

/** Taylor-Hood geometry: vertex coordinates, vertex ids and DOF table. */
struct DeviceTHGeometry
{
  Index             num_elems = 0, ndof = 0;
  DeviceVector      xv, yv;
  DeviceIndexVector vconn3, dofs15;
};

inline void assemble(const TaylorHoodNS& op, const DeviceTHGeometry& geom, const DeviceAssemblyMap& map,
                     const DeviceVector& state, DeviceVector& res, DeviceCsrMatrix* jac,
                     HipContext& ctx)
{
  checkAssembly(map, &state, res, jac);
  require(geom.num_elems == map.info.num_elems, "Geometry and AssemblyMap have different element counts");
  // over the blocks' round tables (fx_th_assemble_rounds; bitwise equal to
  // fx_th_assemble over the lane map)
  check(fx_th_assemble_rounds(geom.ndof, geom.num_elems, map.num_blocks, map.max_block_nnz,
                              geom.xv.data(), geom.yv.data(), geom.vconn3.data(),
                              geom.dofs15.data(), state.data(), op.nu, op.nu_e,
                              map.blk_row.data(), map.blk_flags.data(), map.blk_lanes.data(),
                              map.rt_off.data(), map.rt_elem.data(), map.rt_meta.data(),
                              map.pattern->rowPtrData(), jac ? jac->valsData() : nullptr,
                              res.data(), ctx.raw()));
}

/** Taylor-Hood adjoint VJP (fx_th_vjp): jt = J^T adj over the row-replaced
 * Newton Jacobian (adj zeroed on the rows of bc, passed through there; bc
 * may be null) and, when g is non-null, g[e] = adj~_e^T dR_e/dnu_e.  The
 * Residual::applyParamJacT contract (femx/state/Residual.hpp:55-59) for a
 * per-element viscosity. */
inline void applyTaylorHoodJacT(const TaylorHoodNS& op, const DeviceTHGeometry& geom,
                                const DeviceAssemblyMap& map, const DeviceVector& state,
                                const DeviceVector& adj, const DeviceIndexVector* bc_mask,
                                DeviceVector& jt, DeviceVector* g, HipContext& ctx)
{
  require(map.elem_inc.size() == 15 * geom.num_elems && geom.num_elems == map.info.num_elems,
          "Taylor-Hood VJP needs the 15x15 map of the geometry");
  require(state.size() == geom.ndof && adj.size() == geom.ndof,
          "Taylor-Hood VJP vector sizes do not match");
  if (jt.size() != geom.ndof) jt.resize(geom.ndof);
  if (g && g->size() != geom.num_elems) g->resize(geom.num_elems);
  if (map.vjp_work.size() != 15 * geom.num_elems) map.vjp_work.resize(15 * geom.num_elems);
  check(fx_th_vjp(geom.ndof, geom.num_elems, geom.xv.data(), geom.yv.data(), geom.vconn3.data(),
                  geom.dofs15.data(), state.data(), op.nu, op.nu_e, adj.data(),
                  bc_mask ? bc_mask->data() : nullptr, map.inc_off.data(), map.elem_inc.data(),
                  jt.data(), g ? g->data() : nullptr, map.vjp_work.data(), ctx.raw()));
}

namespace detail
{
/** The GLS node plan (fx_map_gls_plan) on the device. */
struct DeviceNavierPlan
{
  DeviceIndexVector node_rows, inc_off, inc_elem, inc_meta, inc_slot, blk_runs, node_tile;
  Index             max_tile = 0;
  // two-phase history VJP workspace (6 doubles per element), sized on first
  // use; history VJPs of one plan must not overlap
  mutable DeviceVector vjp_work;
};
} // namespace detail

/** Mesh of femx's Navier-Stokes model: node coordinates, 3 node ids per
 * element, and the map's element DOF table (9 per element, NavierOperator
 * local order u0 v0 u1 v1 u2 v2 p0 p1 p2), with the node plan built from the
 * host map. */
struct DeviceNavierGeometry
{
  Index                    num_nodes = 0, num_elems = 0, num_states = 0;
  DeviceVector             x, y;
  DeviceIndexVector        conn3, dofs9;
  detail::DeviceNavierPlan plan;
};

inline void makeNavierGeometry(const std::vector<Real>& x, const std::vector<Real>& y,
                               const std::vector<Index>& conn3, const HostAssemblyMap& map,
                               DeviceNavierGeometry& out, HipContext& ctx)
{
  require(x.size() == y.size() && conn3.size() % 3 == 0,
          "Navier geometry arrays have inconsistent sizes");
  const Index N = static_cast<Index>(x.size()), E = static_cast<Index>(conn3.size() / 3);
  require(map.numElems() == E, "Geometry and AssemblyMap have different element counts");
  out.num_nodes  = N;
  out.num_elems  = E;
  out.num_states = map.numStates();
  out.x.assign(x, ctx);
  out.y.assign(y, ctx);
  out.conn3.assign(conn3, ctx);
  // the map's element DOF table, read back through the row incidence export
  // (incidence t is element DOF slot inc_pos[t] of row r)
  std::vector<Index> d9(9 * static_cast<std::size_t>(E));
  std::vector<Index> inc_off(map.numRes() + 1), inc_pos(map.info().num_inc);
  check(fx_map_export_rowinc(map.handle(), inc_off.data(), inc_pos.data(), nullptr));
  for (Index r = 0; r < map.numRes(); ++r)
    for (Index t = inc_off[r]; t < inc_off[r + 1]; ++t) d9[inc_pos[t]] = r;
  out.dofs9.assign(d9, ctx);
  const std::size_t nb = (N + FX_GLS_BLOCK_NODES - 1) / FX_GLS_BLOCK_NODES;
  std::vector<Index> nr(3 * std::size_t(N)), io(N + 1), ie(3 * std::size_t(E)),
      im(24 * std::size_t(E)), is(3 * std::size_t(E)), br(4 * FX_GLS_MAX_RUNS * nb),
      nt(3 * std::size_t(N));
  Index mx = 0;
  check(fx_map_gls_plan(map.handle(), N, conn3.data(), nr.data(), io.data(), ie.data(), im.data(),
                        is.data(), br.data(), nt.data(), &mx));
  auto& p = out.plan;
  p.node_rows.assign(nr, ctx);
  p.inc_off.assign(io, ctx);
  p.inc_elem.assign(ie, ctx);
  p.inc_meta.assign(im, ctx);
  p.inc_slot.assign(is, ctx);
  p.blk_runs.assign(br, ctx);
  p.node_tile.assign(nt, ctx);
  p.max_tile = mx;
}

} // namespace assembly

namespace state
{
/** Variable block of a time-residual Jacobian (femx/state/TimeResidual.hpp:
 * 94-165): the next state, one history lag, or the parameters. */
class VariableBlock
{
public:
  enum class Kind
  {
    HistoryState,
    NextState,
    Param
  };
  constexpr VariableBlock(Kind k = Kind::HistoryState, Index lag = 0) : kind_(k), lag_(lag) {}
  static constexpr VariableBlock hist(Index lag) { return {Kind::HistoryState, lag}; }
  constexpr bool  isNextState() const { return kind_ == Kind::NextState; }
  constexpr bool  isHistoryState() const { return kind_ == Kind::HistoryState; }
  constexpr bool  isParam() const { return kind_ == Kind::Param; }
  constexpr Index historyLag() const { return lag_; }
  static const VariableBlock NextState;
  static const VariableBlock Param;

private:
  Kind  kind_;
  Index lag_;
};
inline constexpr VariableBlock VariableBlock::NextState{VariableBlock::Kind::NextState};
inline constexpr VariableBlock VariableBlock::Param{VariableBlock::Kind::Param};
} // namespace state

namespace model::ns
{
/** femx::model::ns::FluidParams (femx/model/ns/Config.hpp:10-14). */
struct FluidParams
{
  Real rho = 1.0;
  Real mu  = 1.0;
};

/** NavierOperator(data, fluid, dt) (femx/model/ns/Components.hpp:400-577):
 * the GLS P1/P1 row operator of one time step; here it refers to the mesh
 * (which replaces femx's NavierData: element values are recomputed on the
 * device) and carries the fluid and the step. */
struct NavierOperator
{
  const assembly::DeviceNavierGeometry* geom = nullptr;
  FluidParams                           fluid;
  Real                                  dt = 0.0;
};

/** applyNavierHistJacT (femx/model/ns/NavierStokesModel.cpp:321-363):
 * out = (dR/dhist_lag)^T adj, hand-derived where femx uses Enzyme. */
inline void applyNavierHistJacT(const NavierOperator& op, Index step, Index num_hist, Index lag,
                                DeviceConstVectorView hist, DeviceConstVectorView nxt,
                                DeviceConstVectorView adj, DeviceVector& out, HipContext& ctx)
{
  const auto& g = *op.geom;
  require(hist.size() == num_hist * g.num_states && nxt.size() == g.num_states
              && adj.size() == g.num_states,
          "Navier history VJP dimensions do not match");
  if (out.size() != g.num_states) out.resize(g.num_states);
  if (g.plan.vjp_work.size() != 6 * g.num_elems) g.plan.vjp_work.resize(6 * g.num_elems);
  check(fx_gls_hist_vjp(g.num_nodes, g.num_states, g.num_elems, g.x.data(), g.y.data(),
                        g.conn3.data(), g.dofs9.data(), g.plan.node_rows.data(), op.fluid.rho,
                        op.fluid.mu, op.dt, step, num_hist, lag, hist.data(), nxt.data(),
                        adj.data(), g.plan.inc_off.data(), g.plan.inc_slot.data(), out.data(),
                        g.plan.vjp_work.data(), ctx.raw()));
}
} // namespace model::ns

namespace assembly
{
/** Time assembly of the GLS operator, femx's signature (host
 * Assembly.hpp:235-297, device CudaAssembly.hpp:366-417): res and, for the
 * next-state block, the CSR Jacobian (jac may be null: assembleResidual);
 * with respect to a history block the Jacobian is zero, as evalRow leaves
 * it.  Outputs replaced. */
inline void assemble(const model::ns::NavierOperator& op, Index step, Index num_hist,
                     state::VariableBlock wrt, const DeviceAssemblyMap& map,
                     DeviceConstVectorView hist, DeviceConstVectorView nxt, DeviceVector& res,
                     DeviceCsrMatrix* jac, HipContext& ctx)
{
  const auto& g = *op.geom;
  require(num_hist > 0 && hist.size() == num_hist * map.info.num_states
              && nxt.size() == map.info.num_states,
          "Time assembly state dimensions do not match AssemblyMap");
  require(!wrt.isParam() && (!wrt.isHistoryState()
                             || (wrt.historyLag() >= 0 && wrt.historyLag() < num_hist)),
          "Time assembly variable block is invalid");
  require(g.num_elems == map.info.num_elems, "Geometry and AssemblyMap have different element counts");
  if (res.size() != map.info.num_res) res.resize(map.info.num_res);
  if (jac)
    require(jac->pattern().layoutId() == map.pattern->layoutId(),
            "Time assembly matrix must use the AssemblyMap CSR layout");
  const bool next = wrt.isNextState();
  check(fx_gls_assemble(g.num_nodes, g.num_states, g.num_elems, map.info.nnz, g.plan.max_tile,
                        g.x.data(), g.y.data(), g.conn3.data(), g.dofs9.data(),
                        g.plan.node_rows.data(), op.fluid.rho, op.fluid.mu, op.dt, step, num_hist,
                        hist.data(), nxt.data(), g.plan.inc_off.data(), g.plan.inc_elem.data(),
                        g.plan.inc_meta.data(), g.plan.blk_runs.data(), g.plan.node_tile.data(),
                        jac && next ? jac->valsData() : nullptr, res.data(), ctx.raw()));
  if (jac && !next) ::femx::detail::fill(jac->valsData(), map.info.nnz, 0.0, ctx.raw());
}

} // namespace assembly

namespace linalg
{

/** MatrixHandler<HipCsrBackend> (femx/linalg/handler/MatrixHandler.hpp:
 * 193-378): CSR products on views, the explicit transpose with its pattern
 * and permutation cached per source layout, copy, zero and finalize. */
class HipMatrixHandler
{
public:
  explicit HipMatrixHandler(HipContext& ctx) : ctx_(ctx) {}

  /** y = alpha A x + beta y (MatrixHandler.hpp:270-275); x and y must not
   * overlap each other or the values (MatrixHandler.cu:182-202). */
  void matvec(const DeviceCsrMatrix& A, DeviceConstVectorView x, DeviceVectorView y,
              Real alpha = 1.0, Real beta = 0.0) const
  {
    checkApply(A, x, y, A.cols(), A.rows(), "CSR matvec");
    if (const auto* off = A.pattern().colOff16(ctx_)) // narrower index stream, same results
      check(fx_csr_matvec16(A.rows(), A.cols(), A.pattern().rowPtrData(), off->data(), A.valsData(),
                            x.data(), y.data(), alpha, beta, ctx_.raw()));
    else
      check(fx_csr_matvec(A.rows(), A.cols(), A.pattern().rowPtrData(), A.pattern().colIndData(),
                          A.valsData(), x.data(), y.data(), alpha, beta, ctx_.raw()));
  }
  /** y = alpha A^T x + beta y (MatrixHandler.hpp:286-291), each column in
   * ascending row order. */
  void matvecT(const DeviceCsrMatrix& A, DeviceConstVectorView x, DeviceVectorView y,
               Real alpha = 1.0, Real beta = 0.0) const
  {
    checkApply(A, x, y, A.rows(), A.cols(), "CSR matvecT");
    const auto& t = A.pattern().transposeMap(ctx_);
    check(fx_csr_matvec_t(A.rows(), A.cols(), t.t_ptr.data(), t.t_row.data(), t.t_perm.data(),
                          A.valsData(), x.data(), y.data(), alpha, beta, ctx_.raw()));
  }
  /** out = A x, out resized (MatrixHandler.hpp:303-306). */
  void matvec(const DeviceCsrMatrix& A, DeviceConstVectorView x, DeviceVector& out) const
  {
    if (out.size() != A.rows()) out.resize(A.rows());
    matvec(A, x, view(out));
  }
  /** out = A^T x, out resized (MatrixHandler.hpp:318-321). */
  void matvecT(const DeviceCsrMatrix& A, DeviceConstVectorView x, DeviceVector& out) const
  {
    if (out.size() != A.cols()) out.resize(A.cols());
    matvecT(A, x, view(out));
  }
  /** dst = src^T (MatrixHandler.hpp:250-262): a changed source layout
   * replaces dst's pattern (the transposed pattern is built once per source
   * pattern and cached with it); values move through the cached
   * permutation, a gather. */
  void transpose(const DeviceCsrMatrix& src, DeviceCsrMatrix& dst) const
  {
    require(&src != &dst && (src.nnz() == 0 || src.valsData() != dst.valsData()),
            "CSR transpose does not support in-place output");
    const auto tp = src.pattern().transposed(ctx_);
    if (!dst.hasPattern() || dst.pattern().layoutId() != tp->layoutId()) dst = DeviceCsrMatrix(tp);
    const auto& t = src.pattern().transposeMap(ctx_);
    check(fx_vec_gather(src.nnz(), t.t_perm.data(), src.valsData(), src.nnz(), dst.valsData(),
                        ctx_.raw()));
  }
  /** Same-layout value copy (MatrixHandler.hpp:230-238). */
  void copy(const DeviceCsrMatrix& src, DeviceCsrMatrix& dst) const
  {
    require(src.pattern().layoutId() == dst.pattern().layoutId(),
            "CsrMatrix copy requires compatible source and destination graphs");
    check(fx_vec_axpby(src.nnz(), 1.0, src.valsData(), 0.0, dst.valsData(), ctx_.raw()));
  }
  void zero(DeviceCsrMatrix& A) const { detail::fill(A.valsData(), A.nnz(), 0.0, ctx_.raw()); }
  /** Nothing to do after device assembly (MatrixHandler.hpp:343-346). */
  void finalize(DeviceCsrMatrix&) const noexcept {}

private:
  static void checkApply(const DeviceCsrMatrix& A, DeviceConstVectorView x, DeviceVectorView y,
                         Index nx, Index ny, const char* what)
  {
    require(x.isValid() && y.isValid() && x.size() == nx && y.size() == ny,
            std::string(what) + " vector size mismatch");
    require(!detail::overlap(x.data(), sizeof(Real) * x.size(), y.data(), sizeof(Real) * y.size())
                && !detail::overlap(y.data(), sizeof(Real) * y.size(), A.valsData(),
                                    sizeof(Real) * A.nnz()),
            "CSR matvec input and output must not alias");
  }
  HipContext& ctx_;
};

/** VectorHandler<HipCsrBackend> (femx/linalg/handler/VectorHandler.hpp:
 * 190-477): axpby, copy, zero, gather/scatter and the deterministic dot. */
class HipVectorHandler
{
public:
  explicit HipVectorHandler(HipContext& ctx) : ctx_(ctx), work_(fx_dot_work_len()) {}
  /** y = a x + b y (VectorHandler.hpp:381-386). */
  void axpby(Real a, DeviceConstVectorView x, Real b, DeviceVectorView y) const
  {
    require(x.isValid() && y.isValid() && x.size() == y.size(), "axpby size mismatch");
    check(fx_vec_axpby(y.size(), a, x.data(), b, y.data(), ctx_.raw()));
  }
  /** dst = src (VectorHandler::copy). */
  void copy(DeviceConstVectorView src, DeviceVectorView dst) const
  {
    require(src.size() == dst.size(), "Vector copy size mismatch");
    check(fx_vec_axpby(dst.size(), 1.0, src.data(), 0.0, dst.data(), ctx_.raw()));
  }
  /** out = src, out resized. */
  void copy(DeviceConstVectorView src, DeviceVector& out) const
  {
    if (out.size() != src.size()) out.resize(src.size());
    copy(src, view(out));
  }
  void zero(DeviceVectorView v) const { detail::fill(v.data(), v.size(), 0.0, ctx_.raw()); }
  /** out resized if needed, then zeroed (VectorHandler.hpp:282-300). */
  void resizeOrZero(DeviceVector& out, Index n) const
  {
    if (out.size() != n) out.resize(n);
    zero(view(out));
  }
  /** dst[i] = src[indices[i]] (VectorHandler.hpp:397-399). */
  void gather(DeviceConstVectorView src, DeviceConstIndexView indices, DeviceVectorView dst) const
  {
    require(indices.size() == dst.size(), "Vector gather size mismatch");
    check(fx_vec_gather(dst.size(), indices.data(), src.data(), src.size(), dst.data(), ctx_.raw()));
  }
  /** dst[indices[i]] = src[i] (VectorHandler.hpp:410-412); indices distinct. */
  void scatter(DeviceConstVectorView src, DeviceConstIndexView indices, DeviceVectorView dst) const
  {
    require(indices.size() == src.size(), "Vector scatter size mismatch");
    check(fx_vec_scatter(src.size(), indices.data(), src.data(), dst.data(), dst.size(), ctx_.raw()));
  }
  /** out[0] = x . y, a fixed-order two-stage reduction (VectorHandler.hpp:
   * 423-425). */
  void dot(DeviceConstVectorView x, DeviceConstVectorView y, DeviceVectorView out)
  {
    require(x.size() == y.size() && out.size() == 1, "dot size mismatch");
    check(fx_vec_dot(x.size(), x.data(), y.data(), work_.data(), out.data(), ctx_.raw()));
  }

private:
  HipContext&  ctx_;
  DeviceVector work_;
};

} // namespace linalg

namespace state
{
/** Sizes of a parameter-dependent stationary residual (femx/state/
 * Residual.hpp:9-14). */
struct Dimensions
{
  Index num_states{0};
  Index num_param{0};
  Index num_res{0};
};

/** Stationary residual contract over device storage (femx's
 * Residual<Backend>, femx/state/Residual.hpp:16-60, for the HIP backend:
 * Vec DeviceVector, Mat DeviceCsrMatrix, Pattern DeviceCsrPattern, Ctx
 * HipContext). */
class Residual
{
public:
  using Vec     = DeviceVector;
  using Mat     = DeviceCsrMatrix;
  using Pattern = DeviceCsrPattern;
  using Ctx     = HipContext;

  virtual ~Residual() = default;
  virtual Dimensions dims() const = 0;
  /** The host pattern used for metadata (boundary maps, transposes). */
  virtual const HostCsrPattern& hostPattern() const = 0;
  /** The state-Jacobian pattern in device storage. */
  virtual const Pattern& pattern() const = 0;
  /** out = R(state, prm). */
  virtual void res(const Vec& state, const Vec& prm, Vec& out, Ctx& ctx) const = 0;
  /** out = dR/dstate at (state, prm). */
  virtual void assembleStateJac(const Vec& state, const Vec& prm, Mat& out, Ctx& ctx) const = 0;
  /** out = (dR/dprm)^T adj. */
  virtual void applyParamJacT(const Vec& state, const Vec& prm, const Vec& adj, Vec& out,
                              Ctx& ctx) const = 0;
};

/** P1 Poisson with a distributed control m_e = log kappa_e and Dirichlet
 * rows replaced (femx's PoissonOpt residual, examples/poisson/PoissonOpt.cpp:
 * 125-161, with the control moved from the boundary into the coefficient):
 *   R_i = (K(exp m) u - F)_i on free rows,  R_r = u_r - g_r on constrained rows.
 * Its state Jacobian is the row-replaced stiffness (diagonal one); the
 * transposed products zero the adjoint on constrained rows first, as
 * ConstrainedTimeResidual::applyJacT does (femx/assembly/
 * ConstrainedTimeResidual.cpp:379-405).  The geometry, maps, load and
 * boundary values must outlive the residual.  Calls on one residual share
 * its scratch vectors: they must not overlap. */
class PoissonLogKappaResidual : public Residual
{
public:
  PoissonLogKappaResidual(const assembly::DeviceGeometry& geom, const assembly::DeviceAssemblyMap& map,
                          const assembly::DeviceBoundaryMap* bc, const DeviceVector* bc_vals,
                          const Real* f)
      : geom_(geom), map_(map), bc_(bc), bc_vals_(bc_vals), f_(f), tmp_(map.info.num_res),
        jt_(map.info.num_states), g_(geom.num_elems)
  {
    require(!bc || (bc_vals && bc_vals->size() == bc->num_bcs),
            "PoissonLogKappaResidual needs one boundary value per constrained row");
    require(geom.num_elems == map.info.num_elems && map.info.max_res == 3 && map.info.max_state == 3,
            "PoissonLogKappaResidual needs the P1 map of the geometry");
  }

  Dimensions dims() const override
  {
    return {map_.info.num_states, geom_.num_elems, map_.info.num_res};
  }
  const HostCsrPattern& hostPattern() const override { return map_.pattern->host(); }
  const Pattern&        pattern() const override { return *map_.pattern; }

  void res(const Vec& state, const Vec& prm, Vec& out, Ctx& ctx) const override
  {
    checkPrm(prm);
    if (out.size() != map_.info.num_res) out.resize(map_.info.num_res);
    assembly::assemble(assembly::PoissonP1{prm.data(), true, f_}, geom_, map_, &state, out, nullptr,
                       ctx);
    if (bc_) assembly::replaceRes(*bc_, state, *bc_vals_, out, ctx);
  }

  void assembleStateJac(const Vec& state, const Vec& prm, Mat& out, Ctx& ctx) const override
  {
    checkPrm(prm);
    assembly::assemble(assembly::PoissonP1{prm.data(), true, f_}, geom_, map_, &state, tmp_, &out,
                       ctx);
    if (bc_) assembly::replaceRows(*bc_, out, 1.0, ctx);
  }

  /** out[e] = ((dR/dm)^T adj)_e (Residual::applyParamJacT, Residual.hpp:55-59). */
  void applyParamJacT(const Vec& state, const Vec& prm, const Vec& adj, Vec& out,
                      Ctx& ctx) const override
  {
    if (out.size() != geom_.num_elems) out.resize(geom_.num_elems);
    vjp(state, prm, adj, jt_, out, ctx);
  }

  /** out = J^T adj over the row-replaced state Jacobian. */
  void applyStateJacT(const Vec& state, const Vec& prm, const Vec& adj, Vec& out, Ctx& ctx) const
  {
    if (out.size() != map_.info.num_states) out.resize(map_.info.num_states);
    vjp(state, prm, adj, out, g_, ctx);
  }

  /** Both products in one launch (fx_p1_vjp). */
  void vjp(const Vec& state, const Vec& prm, const Vec& adj, Vec& jt, Vec& g, Ctx& ctx) const
  {
    checkPrm(prm);
    require(state.size() == map_.info.num_states && adj.size() == map_.info.num_res
                && jt.size() == map_.info.num_states && g.size() == geom_.num_elems,
            "PoissonLogKappaResidual VJP vector sizes do not match");
    check(fx_p1_vjp(geom_.num_nodes, geom_.num_elems, geom_.x.data(), geom_.y.data(),
                    geom_.conn.data(), prm.data(), state.data(), adj.data(),
                    bc_ ? bc_->bc_mask.data() : nullptr, map_.inc_off.data(), map_.inc.data(),
                    jt.data(), g.data(), ctx.raw()));
  }

private:
  void checkPrm(const Vec& prm) const
  {
    require(prm.size() == geom_.num_elems, "PoissonLogKappaResidual control size must be num_elems");
  }

  const assembly::DeviceGeometry&    geom_;
  const assembly::DeviceAssemblyMap& map_;
  const assembly::DeviceBoundaryMap* bc_;
  const DeviceVector*                bc_vals_;
  const Real*                        f_;
  mutable DeviceVector               tmp_, jt_, g_;
};

/** Steady Taylor-Hood Navier-Stokes with the per-element viscosity nu_e as
 * parameter and Dirichlet rows replaced, the Residual contract of the
 * north-star NS inverse problem:
 *   R = R_NS(u; nu_e) on free rows,  R_r = u_r - g_r on constrained rows;
 * the state Jacobian is the row-replaced Newton Jacobian, and applyParamJacT
 * gives g_e = adj~_e^T dR_e/dnu_e with the adjoint zeroed on constrained rows
 * (fx_th_vjp).  Inputs must outlive the residual; calls must not overlap. */
class NavierStokesResidual : public Residual
{
public:
  NavierStokesResidual(const assembly::DeviceTHGeometry& geom, const assembly::DeviceAssemblyMap& map,
                       const assembly::DeviceBoundaryMap* bc, const DeviceVector* bc_vals)
      : geom_(geom), map_(map), bc_(bc), bc_vals_(bc_vals), tmp_(map.info.num_res),
        jt_(geom.ndof)
  {
    require(!bc || (bc_vals && bc_vals->size() == bc->num_bcs),
            "NavierStokesResidual needs one boundary value per constrained row");
    require(geom.num_elems == map.info.num_elems && map.info.num_res == geom.ndof
                && map.elem_inc.size() == 15 * geom.num_elems,
            "NavierStokesResidual needs the Taylor-Hood map of the geometry");
  }

  Dimensions dims() const override { return {geom_.ndof, geom_.num_elems, geom_.ndof}; }
  const HostCsrPattern& hostPattern() const override { return map_.pattern->host(); }
  const Pattern&        pattern() const override { return *map_.pattern; }

  void res(const Vec& state, const Vec& nu_e, Vec& out, Ctx& ctx) const override
  {
    checkPrm(nu_e);
    if (out.size() != geom_.ndof) out.resize(geom_.ndof);
    assembly::assemble(assembly::TaylorHoodNS{0.0, nu_e.data()}, geom_, map_, state, out, nullptr,
                       ctx);
    if (bc_) assembly::replaceRes(*bc_, state, *bc_vals_, out, ctx);
  }

  void assembleStateJac(const Vec& state, const Vec& nu_e, Mat& out, Ctx& ctx) const override
  {
    checkPrm(nu_e);
    assembly::assemble(assembly::TaylorHoodNS{0.0, nu_e.data()}, geom_, map_, state, tmp_, &out,
                       ctx);
    if (bc_) assembly::replaceRows(*bc_, out, 1.0, ctx);
  }

  void applyParamJacT(const Vec& state, const Vec& nu_e, const Vec& adj, Vec& out,
                      Ctx& ctx) const override
  {
    checkPrm(nu_e);
    assembly::applyTaylorHoodJacT(assembly::TaylorHoodNS{0.0, nu_e.data()}, geom_, map_, state, adj,
                                  bc_ ? &bc_->bc_mask : nullptr, jt_, &out, ctx);
  }

  /** out = J^T adj over the row-replaced Newton Jacobian. */
  void applyStateJacT(const Vec& state, const Vec& nu_e, const Vec& adj, Vec& out, Ctx& ctx) const
  {
    checkPrm(nu_e);
    assembly::applyTaylorHoodJacT(assembly::TaylorHoodNS{0.0, nu_e.data()}, geom_, map_, state, adj,
                                  bc_ ? &bc_->bc_mask : nullptr, out, nullptr, ctx);
  }

private:
  void checkPrm(const Vec& nu_e) const
  {
    require(nu_e.size() == geom_.num_elems, "NavierStokesResidual viscosity size must be num_elems");
  }

  const assembly::DeviceTHGeometry&  geom_;
  const assembly::DeviceAssemblyMap& map_;
  const assembly::DeviceBoundaryMap* bc_;
  const DeviceVector*                bc_vals_;
  mutable DeviceVector               tmp_, jt_;
};
} // namespace state

} // namespace femx
