// Host-side builders of the femx_amd C-ABI: structured meshes, the assembly
// map (CSR pattern, femx's jac_map and the row-incidence tables of the
// row-gather kernels), the boundary map and the CSR transpose permutation.
//
// The map builder produces femx's exact arrays (femx/assembly/
// AssemblyMap.cpp:54-176) without femx's global COO sort: it buckets the
// local rows by global row (stable in element order), then sorts and
// deduplicates each row's columns on its own.  Rows are independent, so the
// per-row passes run on all host threads.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fx_common.hpp"

using Index = fx_index;

namespace fx
{
namespace
{
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

namespace
{

// Run fn(begin, end) over [0, n) on up to hardware_concurrency() threads.
template <class F>
void parallel_ranges(std::int64_t n, F&& fn)
{
  unsigned nt = std::max(1u, std::thread::hardware_concurrency());
  nt          = static_cast<unsigned>(std::min<std::int64_t>(nt, std::max<std::int64_t>(1, n / 4096)));
  if (nt <= 1)
  {
    fn(std::int64_t(0), n);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(nt);
  for (unsigned t = 0; t < nt; ++t)
  {
    const std::int64_t b = n * t / nt, e = n * (t + 1) / nt;
    pool.emplace_back([&fn, b, e] { fn(b, e); });
  }
  for (auto& th : pool) th.join();
}

} // namespace
} // namespace fx

struct fx_map
{
  Index num_elems = 0, num_res = 0, num_states = 0;
  Index max_res = 0, max_state = 0, max_jac = 0, max_row_len = 0;
  std::vector<Index> res_off, res_dofs, st_off, st_dofs, jac_off;
  std::vector<Index> row_ptr, col_ind, jac_map;
  std::vector<Index> inc_off, inc_elem, inc_local; // row incidence, ascending element per row
};

struct fx_bcmap
{
  Index rows = 0;
  std::vector<Index> bc_rows, diag, col_off, col_ent, col_rows, mask;
  std::vector<Index> elim_rows, elim_off, elim_k, elim_ib;
};

extern "C" {

const char* fx_last_error(void) { return fx::g_last_error.c_str(); }

int fx_version(void) { return 1; }

// ---------------------------------------------------------------------------
// Structured meshes (conventions stated in include/femx_amd.h).  Coordinates
// are x_min + i*dx as femx's Mesh::makeStructuredQuad computes them
// (femx/fem/Mesh.cpp:17-28).

int fx_mesh_p1(Index n, double* x, double* y, Index* conn)
{
  return fx::guard([&] {
    fx::require(n > 0 && x && y && conn, FX_ERR_ARG, "fx_mesh_p1: invalid arguments");
    fx::require(static_cast<std::int64_t>(n + 1) * (n + 1) < std::numeric_limits<Index>::max()
                    && 6LL * n * n < std::numeric_limits<Index>::max(),
                FX_ERR_OVERFLOW, "fx_mesh_p1: mesh exceeds the int32 index range");
    const double dx = 1.0 / static_cast<double>(n);
    const Index  np = n + 1;
    for (Index j = 0; j <= n; ++j)
      for (Index i = 0; i <= n; ++i)
      {
        x[j * np + i] = 0.0 + static_cast<double>(i) * dx;
        y[j * np + i] = 0.0 + static_cast<double>(j) * dx;
      }
    for (Index j = 0; j < n; ++j)
      for (Index i = 0; i < n; ++i)
      {
        const Index a = j * np + i, b = a + 1, d = a + np, c = d + 1;
        Index*      e = conn + 6 * (static_cast<std::int64_t>(j) * n + i);
        e[0] = a, e[1] = b, e[2] = c;
        e[3] = a, e[4] = c, e[5] = d;
      }
  });
}

int fx_mesh_p2(Index n, double* x, double* y, Index* conn6, Index* vconn3)
{
  return fx::guard([&] {
    fx::require(n > 0 && x && y && conn6 && vconn3, FX_ERR_ARG, "fx_mesh_p2: invalid arguments");
    fx::require(static_cast<std::int64_t>(2 * n + 1) * (2 * n + 1) * 2 + (n + 1) * (n + 1)
                        < std::numeric_limits<Index>::max()
                    && 12LL * n * n < std::numeric_limits<Index>::max(),
                FX_ERR_OVERFLOW, "fx_mesh_p2: mesh exceeds the int32 index range");
    const Index  m = 2 * n, mp = m + 1, np = n + 1;
    const double dx = 1.0 / static_cast<double>(m);
    for (Index J = 0; J <= m; ++J)
      for (Index I = 0; I <= m; ++I)
      {
        x[J * mp + I] = 0.0 + static_cast<double>(I) * dx;
        y[J * mp + I] = 0.0 + static_cast<double>(J) * dx;
      }
    auto L = [mp](Index I, Index J) { return J * mp + I; };
    for (Index j = 0; j < n; ++j)
      for (Index i = 0; i < n; ++i)
      {
        const Index        I = 2 * i, J = 2 * j;
        const std::int64_t c = static_cast<std::int64_t>(j) * n + i;
        Index*             e = conn6 + 12 * c;
        e[0] = L(I, J), e[1] = L(I + 2, J), e[2] = L(I + 2, J + 2);
        e[3] = L(I + 1, J), e[4] = L(I + 2, J + 1), e[5] = L(I + 1, J + 1);
        e[6] = L(I, J), e[7] = L(I + 2, J + 2), e[8] = L(I, J + 2);
        e[9] = L(I + 1, J + 1), e[10] = L(I + 1, J + 2), e[11] = L(I, J + 1);
        Index*      v = vconn3 + 6 * c;
        const Index a = j * np + i, b = a + 1, d = a + np, cc = d + 1;
        v[0] = a, v[1] = b, v[2] = cc;
        v[3] = a, v[4] = cc, v[5] = d;
      }
  });
}

int fx_th_dofs(Index n, const Index* conn6, const Index* vconn3, Index* dofs15)
{
  return fx::guard([&] {
    fx::require(n > 0 && conn6 && vconn3 && dofs15, FX_ERR_ARG, "fx_th_dofs: invalid arguments");
    const Index        nv = (2 * n + 1) * (2 * n + 1);
    const std::int64_t E  = 2LL * n * n;
    for (std::int64_t e = 0; e < E; ++e)
    {
      Index* d = dofs15 + 15 * e;
      for (int a = 0; a < 6; ++a)
        for (int c = 0; c < 2; ++c) d[2 * a + c] = 2 * conn6[6 * e + a] + c;
      for (int b = 0; b < 3; ++b) d[12 + b] = 2 * nv + vconn3[3 * e + b];
    }
  });
}

int fx_th_mesh(Index num_vertices, Index num_elems, const Index* conn3, Index* num_p2_nodes,
               Index* conn6, Index* edge_ends, Index* dofs15)
{
  return fx::guard([&] {
    fx::require(num_vertices >= 0 && num_elems >= 0 && num_p2_nodes
                    && (num_elems == 0 || conn3),
                FX_ERR_ARG, "fx_th_mesh: invalid arguments");
    // edge (a, b), a < b, -> its node id; first appearance numbers it
    std::unordered_map<std::uint64_t, Index> edge_id;
    std::vector<Index>                       uses, ends;
    edge_id.reserve(static_cast<std::size_t>(num_elems) * 2);
    std::vector<Index> c6(6 * static_cast<std::size_t>(num_elems));
    for (Index e = 0; e < num_elems; ++e)
    {
      const Index* v = conn3 + 3 * static_cast<std::int64_t>(e);
      for (int k = 0; k < 3; ++k)
        fx::require(v[k] >= 0 && v[k] < num_vertices, FX_ERR_ARG,
                    "fx_th_mesh: vertex index out of range");
      fx::require(v[0] != v[1] && v[1] != v[2] && v[2] != v[0], FX_ERR_ARG,
                  "fx_th_mesh: element with a repeated vertex");
      for (int k = 0; k < 3; ++k) c6[6 * std::size_t(e) + k] = v[k];
      for (int k = 0; k < 3; ++k)
      {
        const Index   a = v[k], b = v[(k + 1) % 3];
        const Index   lo = std::min(a, b), hi = std::max(a, b);
        const auto    key = (static_cast<std::uint64_t>(lo) << 32) | static_cast<std::uint32_t>(hi);
        auto          it  = edge_id.find(key);
        Index         id;
        if (it == edge_id.end())
        {
          fx::require(static_cast<std::int64_t>(num_vertices) + uses.size()
                          < std::numeric_limits<Index>::max() / 2,
                      FX_ERR_OVERFLOW, "fx_th_mesh: P2 node count exceeds int32");
          id = num_vertices + static_cast<Index>(uses.size());
          edge_id.emplace(key, id);
          uses.push_back(1);
          ends.push_back(a);
          ends.push_back(b);
        }
        else
        {
          id = it->second;
          fx::require(++uses[id - num_vertices] <= 2, FX_ERR_UNSUPPORTED,
                      "fx_th_mesh: an edge is shared by more than two elements");
        }
        c6[6 * std::size_t(e) + 3 + k] = id;
      }
    }
    const Index np2 = num_vertices + static_cast<Index>(uses.size());
    fx::require(2LL * np2 + num_vertices < std::numeric_limits<Index>::max(), FX_ERR_OVERFLOW,
                "fx_th_mesh: Taylor-Hood DOF count exceeds int32");
    *num_p2_nodes = np2;
    if (conn6) std::copy(c6.begin(), c6.end(), conn6);
    if (edge_ends) std::copy(ends.begin(), ends.end(), edge_ends);
    if (dofs15)
      for (Index e = 0; e < num_elems; ++e)
      {
        Index* d = dofs15 + 15 * static_cast<std::int64_t>(e);
        for (int a = 0; a < 6; ++a)
          for (int c = 0; c < 2; ++c) d[2 * a + c] = 2 * c6[6 * std::size_t(e) + a] + c;
        for (int b = 0; b < 3; ++b) d[12 + b] = 2 * np2 + c6[6 * std::size_t(e) + b];
      }
  });
}

// ---------------------------------------------------------------------------
// Assembly map.  Validation and messages follow femx's makeAssemblyMap
// (femx/assembly/AssemblyMap.cpp:20-100).

int fx_map_create(Index num_res, Index num_states, Index num_elems, const Index* res_offsets,
                  const Index* res_dofs, const Index* state_offsets, const Index* state_dofs,
                  fx_map** out)
{
  return fx::guard([&] {
    fx::require(out != nullptr, FX_ERR_ARG, "fx_map_create: null output handle");
    *out = nullptr;
    fx::require(num_res >= 0 && num_states >= 0,
                FX_ERR_ARG, "AssemblyMap global dimensions must be non-negative");
    fx::require(num_elems >= 0 && res_offsets && state_offsets
                    && (num_elems == 0 || (res_dofs && state_dofs)),
                FX_ERR_ARG, "fx_map_create: invalid element tables");
    auto m        = std::make_unique<fx_map>();
    m->num_elems  = num_elems;
    m->num_res    = num_res;
    m->num_states = num_states;
    m->res_off.assign(res_offsets, res_offsets + num_elems + 1);
    m->st_off.assign(state_offsets, state_offsets + num_elems + 1);
    fx::require(m->res_off[0] == 0 && m->st_off[0] == 0, FX_ERR_ARG,
                "AssemblyMap element offsets must start at zero");
    m->jac_off.assign(num_elems + 1, 0);
    for (Index e = 0; e < num_elems; ++e)
    {
      const Index nr = m->res_off[e + 1] - m->res_off[e];
      const Index nc = m->st_off[e + 1] - m->st_off[e];
      fx::require(nr >= 0 && nc >= 0, FX_ERR_ARG,
                  "AssemblyMap local dimensions must be non-negative");
      const std::int64_t nloc = static_cast<std::int64_t>(nr) * nc;
      fx::require(nloc <= std::numeric_limits<Index>::max(), FX_ERR_OVERFLOW,
                  "AssemblyMap local Jacobian is too large");
      const std::int64_t tot = static_cast<std::int64_t>(m->jac_off[e]) + nloc;
      fx::require(tot <= std::numeric_limits<Index>::max(), FX_ERR_OVERFLOW,
                  "AssemblyMap flattened data exceeds the Index range");
      m->jac_off[e + 1] = static_cast<Index>(tot);
      m->max_res        = std::max(m->max_res, nr);
      m->max_state      = std::max(m->max_state, nc);
      m->max_jac        = std::max(m->max_jac, static_cast<Index>(nloc));
    }
    const Index nres_tot = m->res_off[num_elems], nst_tot = m->st_off[num_elems];
    m->res_dofs.assign(res_dofs, res_dofs + nres_tot);
    m->st_dofs.assign(state_dofs, state_dofs + nst_tot);
    for (Index d : m->res_dofs)
      fx::require(d >= 0 && d < num_res, FX_ERR_ARG, "AssemblyMap residual DOF is out of range");
    for (Index d : m->st_dofs)
      fx::require(d >= 0 && d < num_states, FX_ERR_ARG, "AssemblyMap state DOF is out of range");

    // Row incidence: bucket (element, local row) by global row, stable in
    // element order.
    m->inc_off.assign(num_res + 1, 0);
    for (Index d : m->res_dofs) ++m->inc_off[d + 1];
    for (Index r = 0; r < num_res; ++r) m->inc_off[r + 1] += m->inc_off[r];
    m->inc_elem.resize(nres_tot);
    m->inc_local.resize(nres_tot);
    {
      std::vector<Index> next(m->inc_off.begin(), m->inc_off.end() - 1);
      for (Index e = 0; e < num_elems; ++e)
        for (Index p = m->res_off[e]; p < m->res_off[e + 1]; ++p)
        {
          const Index t  = next[m->res_dofs[p]]++;
          m->inc_elem[t]  = e;
          m->inc_local[t] = p - m->res_off[e];
        }
    }

    // Per-row column sets: pass 1 counts, pass 2 fills col_ind and jac_map.
    std::vector<Index> row_len(num_res, 0);
    auto row_cols = [&](Index r, std::vector<Index>& cols) {
      cols.clear();
      for (Index t = m->inc_off[r]; t < m->inc_off[r + 1]; ++t)
      {
        const Index e = m->inc_elem[t];
        cols.insert(cols.end(), m->st_dofs.begin() + m->st_off[e],
                    m->st_dofs.begin() + m->st_off[e + 1]);
      }
      std::sort(cols.begin(), cols.end());
      cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    };
    fx::parallel_ranges(num_res, [&](std::int64_t b, std::int64_t e) {
      std::vector<Index> cols;
      for (std::int64_t r = b; r < e; ++r)
      {
        row_cols(static_cast<Index>(r), cols);
        row_len[r] = static_cast<Index>(cols.size());
      }
    });
    m->row_ptr.assign(num_res + 1, 0);
    for (Index r = 0; r < num_res; ++r)
    {
      m->row_ptr[r + 1] = m->row_ptr[r] + row_len[r];
      m->max_row_len    = std::max(m->max_row_len, row_len[r]);
    }
    m->col_ind.resize(m->row_ptr[num_res]);
    m->jac_map.resize(m->jac_off[num_elems]);
    fx::parallel_ranges(num_res, [&](std::int64_t b, std::int64_t e) {
      std::vector<Index> cols;
      for (std::int64_t rr = b; rr < e; ++rr)
      {
        const Index r = static_cast<Index>(rr);
        row_cols(r, cols);
        const Index base = m->row_ptr[r];
        std::copy(cols.begin(), cols.end(), m->col_ind.begin() + base);
        for (Index t = m->inc_off[r]; t < m->inc_off[r + 1]; ++t)
        {
          const Index el = m->inc_elem[t], li = m->inc_local[t];
          const Index nc = m->st_off[el + 1] - m->st_off[el];
          Index*      jm = m->jac_map.data() + m->jac_off[el] + li * nc;
          for (Index lj = 0; lj < nc; ++lj)
          {
            const Index c = m->st_dofs[m->st_off[el] + lj];
            jm[lj] = base + static_cast<Index>(std::lower_bound(cols.begin(), cols.end(), c)
                                               - cols.begin());
          }
        }
      }
    });
    *out = m.release();
  });
}

int fx_map_info_get(const fx_map* m, fx_map_info* info)
{
  return fx::guard([&] {
    fx::require(m && info, FX_ERR_ARG, "fx_map_info_get: null argument");
    info->num_elems   = m->num_elems;
    info->num_res     = m->num_res;
    info->num_states  = m->num_states;
    info->nnz         = m->row_ptr.back();
    info->jac_len     = m->jac_off.back();
    info->max_res     = m->max_res;
    info->max_state   = m->max_state;
    info->max_jac     = m->max_jac;
    info->max_row_len = m->max_row_len;
    info->num_inc     = static_cast<Index>(m->inc_elem.size());
  });
}

int fx_map_export_pattern(const fx_map* m, Index* row_ptr, Index* col_ind)
{
  return fx::guard([&] {
    fx::require(m && row_ptr && (col_ind || m->col_ind.empty()), FX_ERR_ARG,
                "fx_map_export_pattern: null argument");
    std::copy(m->row_ptr.begin(), m->row_ptr.end(), row_ptr);
    std::copy(m->col_ind.begin(), m->col_ind.end(), col_ind);
  });
}

int fx_map_export_jac(const fx_map* m, Index* jac_offsets, Index* jac_map)
{
  return fx::guard([&] {
    fx::require(m && jac_offsets && (jac_map || m->jac_map.empty()), FX_ERR_ARG,
                "fx_map_export_jac: null argument");
    std::copy(m->jac_off.begin(), m->jac_off.end(), jac_offsets);
    std::copy(m->jac_map.begin(), m->jac_map.end(), jac_map);
  });
}

int fx_map_export_rowinc3(const fx_map* m, Index* inc_off, Index* inc)
{
  return fx::guard([&] {
    fx::require(m && inc_off && (inc || m->inc_elem.empty()), FX_ERR_ARG,
                "fx_map_export_rowinc3: null argument");
    for (Index e = 0; e < m->num_elems; ++e)
      fx::require(m->res_off[e + 1] - m->res_off[e] == 3 && m->st_off[e + 1] - m->st_off[e] == 3,
                  FX_ERR_UNSUPPORTED, "fx_map_export_rowinc3: every element needs 3x3 local DOFs");
    fx::require(m->max_row_len <= 255, FX_ERR_UNSUPPORTED,
                "fx_map_export_rowinc3: CSR rows longer than 255 entries");
    std::copy(m->inc_off.begin(), m->inc_off.end(), inc_off);
    for (Index r = 0; r < m->num_res; ++r)
      for (Index t = m->inc_off[r]; t < m->inc_off[r + 1]; ++t)
      {
        const Index  e = m->inc_elem[t], li = m->inc_local[t];
        const Index* jm   = m->jac_map.data() + m->jac_off[e] + 3 * li;
        const Index  base = m->row_ptr[r];
        std::uint32_t meta = static_cast<std::uint32_t>(li);
        for (int j = 0; j < 3; ++j)
          meta |= static_cast<std::uint32_t>(jm[j] - base) << (8 * (j + 1));
        inc[2 * t]     = e;
        inc[2 * t + 1] = static_cast<Index>(meta);
      }
  });
}


// Generic row incidence: inc_off[num_res+1] and inc_pos[num_inc], the
// position res_offsets[e] + li of each (element, local row) in the flattened
// element DOF table, rows in ascending element order.
int fx_map_export_rowinc(const fx_map* m, Index* inc_off, Index* inc_pos, Index* elem_inc)
{
  return fx::guard([&] {
    fx::require(m && inc_off, FX_ERR_ARG, "fx_map_export_rowinc: null argument");
    std::copy(m->inc_off.begin(), m->inc_off.end(), inc_off);
    for (std::size_t t = 0; t < m->inc_elem.size(); ++t)
    {
      const Index p = m->res_off[m->inc_elem[t]] + m->inc_local[t];
      if (inc_pos) inc_pos[t] = p;
      if (elem_inc) elem_inc[p] = static_cast<Index>(t);
    }
  });
}

// Row incidence for 15x15 elements (Taylor-Hood): inc_elem[num_inc] and
// inc_meta[4*num_inc] (16 bytes per incidence: positions of the element's 15
// columns within the row, from row_ptr[row], then the local row).
int fx_map_export_rowinc15(const fx_map* m, Index* inc_off, Index* inc_elem, Index* inc_meta)
{
  return fx::guard([&] {
    fx::require(m && inc_off && (m->inc_elem.empty() || (inc_elem && inc_meta)), FX_ERR_ARG,
                "fx_map_export_rowinc15: null argument");
    for (Index e = 0; e < m->num_elems; ++e)
      fx::require(m->res_off[e + 1] - m->res_off[e] == 15 && m->st_off[e + 1] - m->st_off[e] == 15,
                  FX_ERR_UNSUPPORTED, "fx_map_export_rowinc15: every element needs 15x15 local DOFs");
    fx::require(m->max_row_len <= 255, FX_ERR_UNSUPPORTED,
                "fx_map_export_rowinc15: CSR rows longer than 255 entries");
    std::copy(m->inc_off.begin(), m->inc_off.end(), inc_off);
    for (Index r = 0; r < m->num_res; ++r)
      for (Index t = m->inc_off[r]; t < m->inc_off[r + 1]; ++t)
      {
        const Index  e = m->inc_elem[t], li = m->inc_local[t];
        const Index* jm   = m->jac_map.data() + m->jac_off[e] + 15 * li;
        std::uint8_t b[16];
        for (int j = 0; j < 15; ++j) b[j] = static_cast<std::uint8_t>(jm[j] - m->row_ptr[r]);
        b[15] = static_cast<std::uint8_t>(li);
        inc_elem[t] = e;
        std::memcpy(inc_meta + 4 * t, b, 16);
      }
  });
}

// Work blocks of the Taylor-Hood row kernels.  Rows are cut into blocks of
// at most `lanes` work items; a work item is a pair of velocity rows (r, r+1)
// when the two rows are interchangeable for the kernel -- same incident
// elements in the same order, local rows li and li+1 with li even and < 12,
// identical column positions -- and a single row otherwise.  Blocks hold
// contiguous rows and at most min(max_nnz, FX_TH_TILE_NNZ_MAX) CSR entries,
// max_nnz = 0 meaning FX_TH_TILE_NNZ_MAX (a block always takes at least one
// item).  blk_row[num_blocks+1] are block row
// starts, blk_flags the item kind (1 = pairs, 0 = single rows);
// max_block_nnz bounds the LDS row tile.  Call with null arrays to get the
// counts.
int fx_map_th_blocks(const fx_map* m, Index lanes, Index max_nnz, Index* num_blocks,
                     Index* max_block_nnz, Index* blk_row, Index* blk_flags)
{
  return fx::guard([&] {
    fx::require(m && num_blocks && max_block_nnz && lanes > 0 && max_nnz >= 0, FX_ERR_ARG,
                "fx_map_th_blocks: invalid arguments");
    auto pairable = [&](Index r) {
      if (r + 1 >= m->num_res) return false;
      const Index a0 = m->inc_off[r], a1 = m->inc_off[r + 1], b1 = m->inc_off[r + 2];
      if (a1 - a0 != b1 - a1 || a1 == a0) return false;
      for (Index k = 0; k < a1 - a0; ++k)
      {
        const Index t = a0 + k, u = a1 + k;
        const Index e = m->inc_elem[t], li = m->inc_local[t];
        if (m->inc_elem[u] != e || m->inc_local[u] != li + 1 || (li & 1) || li >= 12) return false;
        if (m->res_off[e + 1] - m->res_off[e] != 15 || m->st_off[e + 1] - m->st_off[e] != 15)
          return false;
        const Index* ja = m->jac_map.data() + m->jac_off[e] + 15 * li;
        const Index* jb = ja + 15;
        for (int j = 0; j < 15; ++j)
          if (ja[j] - m->row_ptr[r] != jb[j] - m->row_ptr[r + 1]) return false;
      }
      return true;
    };
    // every block's tile fits the kernels' LDS (a single item always does:
    // two rows of at most 255 entries)
    const Index cap = (max_nnz == 0 || max_nnz > FX_TH_TILE_NNZ_MAX) ? FX_TH_TILE_NNZ_MAX : max_nnz;
    std::vector<Index> starts, flags;
    Index              r = 0, mx = 0;
    while (r < m->num_res)
    {
      const bool  pair  = pairable(r);
      const Index start = r;
      Index       items = 0;
      while (r < m->num_res && items < lanes && pairable(r) == pair)
      {
        const Index next = r + (pair ? 2 : 1);
        if (items > 0 && m->row_ptr[next] - m->row_ptr[start] > cap) break;
        r = next;
        ++items;
      }
      starts.push_back(start);
      flags.push_back(pair ? 1 : 0);
      mx = std::max(mx, m->row_ptr[r] - m->row_ptr[start]);
    }
    starts.push_back(m->num_res);
    *num_blocks    = static_cast<Index>(flags.size());
    *max_block_nnz = mx;
    if (blk_row) std::copy(starts.begin(), starts.end(), blk_row);
    if (blk_flags) std::copy(flags.begin(), flags.end(), blk_flags);
  });
}

int fx_map_th_lanes(const fx_map* m, Index num_blocks, const Index* blk_row, const Index* blk_flags,
                    Index* blk_lanes)
{
  return fx::guard([&] {
    fx::require(m && num_blocks >= 0 && (num_blocks == 0 || (blk_row && blk_flags && blk_lanes)),
                FX_ERR_ARG, "fx_map_th_lanes: invalid arguments");
    constexpr int L = 64;
    std::vector<Index> deg;
    for (Index b = 0; b < num_blocks; ++b)
    {
      const Index rs = blk_row[b], re = blk_row[b + 1], step = blk_flags[b] ? 2 : 1;
      fx::require(rs >= 0 && re > rs && re <= m->num_res && (re - rs) % step == 0, FX_ERR_ARG,
                  "fx_map_th_lanes: blk_row is not a block partition of the rows");
      deg.clear();
      for (Index r = rs; r < re; r += step) deg.push_back(m->inc_off[r + 1] - m->inc_off[r]);
      fx::require(static_cast<Index>(deg.size()) <= L, FX_ERR_UNSUPPORTED,
                  "fx_map_th_lanes: a block holds more than 64 items");
      Index dmax = 1;
      for (Index d : deg) dmax = std::max(dmax, d);
      // fewest rounds whose lane demand fits the wave (R = dmax always fits)
      Index R = 1;
      for (;; ++R)
      {
        Index need = 0;
        for (Index d : deg) need += std::max<Index>(1, (d + R - 1) / R);
        if (need <= L || R >= dmax) break;
      }
      Index* out  = blk_lanes + static_cast<std::int64_t>(L) * b;
      Index  lane = 0;
      for (Index i = 0; i < static_cast<Index>(deg.size()); ++i)
      {
        const Index k = std::max<Index>(1, (deg[i] + R - 1) / R);
        for (Index j = 0; j < k; ++j) out[lane++] = i | (j << 8) | (k << 16);
      }
      for (; lane < L; ++lane) out[lane] = -1;
    }
  });
}

// Round tables of the Taylor-Hood assembly.  Block b runs R_b rounds, the
// most incidences any of its lanes takes under the lane map; the incidence
// that lane l takes in round m is entry rt_off[b] + 64 m + l of rt_elem /
// rt_meta (four words, as fx_map_export_rowinc15 writes them).  An entry past
// the lane's last incidence, or on a lane the map leaves idle, repeats a
// valid incidence of the block -- the lane's last, else the block's first --
// with bit 31 of its last meta word set: it is loaded, never added.  A block
// whose rows have no incidence gets one such round (filled with incidence 0)
// when the mesh has any incidence, so that a kernel walking several blocks
// always finds entries; in a mesh without incidences every R_b is 0.
int fx_map_th_rounds(Index num_rows, Index num_blocks, const Index* blk_row,
                     const Index* blk_flags, const Index* blk_lanes, const Index* inc_off,
                     const Index* inc_elem, const Index* inc_meta, Index* rt_off, Index* rt_elem,
                     Index* rt_meta)
{
  return fx::guard([&] {
    fx::require(num_rows >= 0 && num_blocks >= 0 && rt_off
                    && (num_blocks == 0 || (blk_row && blk_flags && blk_lanes && inc_off))
                    && (!rt_elem || (rt_meta && inc_elem && inc_meta)),
                FX_ERR_ARG, "fx_map_th_rounds: invalid arguments");
    fx::require(num_blocks == 0 || (blk_row[0] >= 0 && blk_row[num_blocks] <= num_rows), FX_ERR_ARG,
                "fx_map_th_rounds: blk_row is not a block partition of the rows");
    for (Index r = 0; r < num_rows; ++r)
      fx::require(inc_off[r] >= 0 && inc_off[r + 1] >= inc_off[r], FX_ERR_ARG,
                  "fx_map_th_rounds: inc_off is not a row offset array");
    constexpr int      L = 64;
    std::vector<Index> first(L), count(L), stride(L);
    std::int64_t       total = 0;
    rt_off[0]                = 0;
    const bool any = num_blocks > 0 && inc_off[blk_row[num_blocks]] > inc_off[blk_row[0]];
    for (Index b = 0; b < num_blocks; ++b)
    {
      const Index rs = blk_row[b], re = blk_row[b + 1], step = blk_flags[b] ? 2 : 1;
      fx::require(re > rs && re <= num_rows && (re - rs) % step == 0, FX_ERR_ARG,
                  "fx_map_th_rounds: blk_row is not a block partition of the rows");
      const Index  nitems = (re - rs) / step;
      const Index* lm     = blk_lanes + static_cast<std::int64_t>(L) * b;
      Index        R = 0, fill = -1;
      for (int l = 0; l < L; ++l)
      {
        count[l] = 0;
        if (lm[l] < 0) continue;
        const Index item = lm[l] & 0xff, role = (lm[l] >> 8) & 0xff,
                    k = std::max<Index>(1, lm[l] >> 16);
        fx::require(item < nitems && role < k, FX_ERR_ARG, "fx_map_th_rounds: lane map out of range");
        const Index r = rs + step * item, t0 = inc_off[r], t1 = inc_off[r + 1];
        // the lane's incidences are t0 + role + m k, m < count
        first[l]  = t0 + role;
        stride[l] = k;
        count[l]  = t0 + role < t1 ? (t1 - t0 - role + k - 1) / k : 0;
        R         = std::max(R, count[l]);
        if (count[l] > 0 && fill < 0) fill = t0 + role;
      }
      if (R == 0 && any) R = 1, fill = inc_off[blk_row[0]];
      total += static_cast<std::int64_t>(L) * R;
      fx::require(total <= std::numeric_limits<Index>::max(), FX_ERR_OVERFLOW,
                  "fx_map_th_rounds: round tables beyond 2^31 entries");
      rt_off[b + 1] = static_cast<Index>(total);
      if (!rt_elem) continue;
      for (Index m = 0; m < R; ++m)
        for (int l = 0; l < L; ++l)
        {
          const Index c = count[l];
          Index       t;
          bool        idle;
          if (c > 0)
            t = first[l] + std::min(m, c - 1) * stride[l], idle = m >= c;
          else
            t = fill, idle = true;
          const std::int64_t q = rt_off[b] + static_cast<std::int64_t>(L) * m + l;
          rt_elem[q]           = inc_elem[t];
          std::memcpy(rt_meta + 4 * q, inc_meta + 4 * static_cast<std::int64_t>(t), 16);
          if (idle)
            rt_meta[4 * q + 3] =
                static_cast<Index>(static_cast<std::uint32_t>(rt_meta[4 * q + 3]) | 0x80000000u);
        }
    }
  });
}

// Node plan of the GLS P1/P1 kernel.  A node's three rows are read off the
// element DOF tables (local rows 2k, 2k+1 and 6+k of local node k); every
// element must agree, and every row must belong to exactly one node.  Blocks
// of FX_GLS_BLOCK_NODES consecutive nodes list their rows as runs of
// consecutive rows, whose CSR ranges are laid end to end in the LDS tile.
int fx_map_gls_plan(const fx_map* m, Index num_nodes, const Index* conn3, Index* node_rows,
                    Index* inc_off, Index* inc_elem, Index* inc_meta, Index* inc_slot,
                    Index* blk_runs, Index* node_tile, Index* max_tile)
{
  return fx::guard([&] {
    fx::require(m && num_nodes >= 0 && node_rows && inc_off && blk_runs && node_tile && max_tile
                    && (m->num_elems == 0 || (conn3 && inc_elem && inc_meta)),
                FX_ERR_ARG, "fx_map_gls_plan: invalid arguments");
    const Index E = m->num_elems, N = num_nodes;
    fx::require(m->num_res == m->num_states, FX_ERR_UNSUPPORTED,
                "fx_map_gls_plan: residual and state spaces differ");
    for (Index e = 0; e < E; ++e)
    {
      fx::require(m->res_off[e + 1] - m->res_off[e] == 9 && m->st_off[e + 1] - m->st_off[e] == 9,
                  FX_ERR_UNSUPPORTED, "fx_map_gls_plan: every element needs 9x9 local DOFs");
      for (int k = 0; k < 9; ++k)
        fx::require(m->res_dofs[m->res_off[e] + k] == m->st_dofs[m->st_off[e] + k],
                    FX_ERR_UNSUPPORTED, "fx_map_gls_plan: residual and state DOF tables differ");
    }
    fx::require(m->max_row_len <= 255, FX_ERR_UNSUPPORTED,
                "fx_map_gls_plan: CSR rows longer than 255 entries");
    std::fill(node_rows, node_rows + 3 * std::size_t(N), Index(-1));
    std::vector<Index> cnt(std::size_t(N) + 1, 0);
    for (Index e = 0; e < E; ++e)
      for (int k = 0; k < 3; ++k)
      {
        const Index v = conn3[3 * e + k];
        fx::require(v >= 0 && v < N, FX_ERR_ARG, "fx_map_gls_plan: node id out of range");
        const Index* d    = m->res_dofs.data() + m->res_off[e];
        const Index  r[3] = {d[2 * k], d[2 * k + 1], d[6 + k]};
        Index*       nr   = node_rows + 3 * std::size_t(v);
        if (nr[0] < 0)
          std::copy(r, r + 3, nr);
        else
          fx::require(nr[0] == r[0] && nr[1] == r[1] && nr[2] == r[2], FX_ERR_UNSUPPORTED,
                      "fx_map_gls_plan: elements disagree on a node's rows");
        ++cnt[v + 1];
      }
    std::vector<Index> owner(m->num_res, -1);
    for (Index n = 0; n < N; ++n)
      for (int k = 0; k < 3 && node_rows[3 * n] >= 0; ++k)
      {
        const Index r = node_rows[3 * n + k];
        fx::require(owner[r] < 0, FX_ERR_UNSUPPORTED,
                    "fx_map_gls_plan: a row belongs to two nodes or twice to one");
        owner[r] = n;
      }
    for (Index r = 0; r < m->num_res; ++r)
      fx::require(owner[r] >= 0, FX_ERR_UNSUPPORTED, "fx_map_gls_plan: a row belongs to no node");
    for (Index n = 0; n < N; ++n) cnt[n + 1] += cnt[n];
    std::copy(cnt.begin(), cnt.end(), inc_off);
    std::vector<Index> li_of(3 * std::size_t(E));
    {
      std::vector<Index> fill(cnt.begin(), cnt.end() - 1);
      for (Index e = 0; e < E; ++e)
        for (int k = 0; k < 3; ++k)
        {
          const Index t = fill[conn3[3 * e + k]]++;
          inc_elem[t]   = e;
          li_of[t]      = k;
          if (inc_slot) inc_slot[3 * e + k] = t;
        }
    }
    std::vector<std::string> errs(1);
    fx::parallel_ranges(N, [&](std::int64_t n0, std::int64_t n1) {
      for (std::int64_t n = n0; n < n1; ++n)
        for (Index t = cnt[n]; t < cnt[n + 1]; ++t)
        {
          const Index  e = inc_elem[t], ln = li_of[t];
          std::uint8_t b[32] = {};
          for (int kk = 0; kk < 3; ++kk)
          {
            const Index  lr   = kk < 2 ? 2 * ln + kk : 6 + ln;
            const Index* jm   = m->jac_map.data() + m->jac_off[e] + 9 * lr;
            const Index  base = m->row_ptr[node_rows[3 * n + kk]];
            for (int j = 0; j < 9; ++j) b[9 * kk + j] = static_cast<std::uint8_t>(jm[j] - base);
          }
          b[27] = static_cast<std::uint8_t>(ln);
          std::memcpy(inc_meta + 8 * std::size_t(t), b, 32);
        }
    });
    const Index nb = (N + FX_GLS_BLOCK_NODES - 1) / FX_GLS_BLOCK_NODES;
    std::fill(blk_runs, blk_runs + std::size_t(nb) * 4 * FX_GLS_MAX_RUNS, Index(0));
    Index mx = 0;
    for (Index b = 0; b < nb; ++b)
    {
      const Index        n0 = b * FX_GLS_BLOCK_NODES, n1 = std::min(N, n0 + FX_GLS_BLOCK_NODES);
      std::vector<Index> rows;
      for (Index n = n0; n < n1; ++n)
        if (node_rows[3 * n] >= 0) rows.insert(rows.end(), node_rows + 3 * n, node_rows + 3 * n + 3);
      std::sort(rows.begin(), rows.end());
      Index* runs = blk_runs + std::size_t(b) * 4 * FX_GLS_MAX_RUNS;
      Index  run_row[FX_GLS_MAX_RUNS + 1];
      int    nr   = 0;
      Index  tile = 0;
      for (std::size_t i = 0; i < rows.size();)
      {
        std::size_t j = i + 1;
        while (j < rows.size() && rows[j] == rows[j - 1] + 1) ++j;
        fx::require(nr < FX_GLS_MAX_RUNS, FX_ERR_UNSUPPORTED,
                    "fx_map_gls_plan: a block's rows need more than FX_GLS_MAX_RUNS runs");
        const Index c0 = m->row_ptr[rows[i]], c1 = m->row_ptr[rows[j - 1] + 1];
        runs[4 * nr]     = c0;
        runs[4 * nr + 1] = c1 - c0;
        runs[4 * nr + 2] = tile;
        run_row[nr]      = rows[i];
        ++nr;
        tile += c1 - c0;
        i = j;
      }
      mx = std::max(mx, tile);
      for (Index n = n0; n < n1; ++n)
        for (int k = 0; k < 3; ++k)
        {
          const Index r = node_rows[3 * n + k];
          int         q = nr - 1;
          while (q > 0 && run_row[q] > r) --q; // the run holding row r
          node_tile[3 * n + k] = r < 0 ? 0 : runs[4 * q + 2] + m->row_ptr[r] - runs[4 * q];
        }
    }
    *max_tile = mx;
  });
}

void fx_map_destroy(fx_map* m) { delete m; }

} // extern "C"

// ---------------------------------------------------------------------------
// Patch plan of the single-pass VJPs (Taylor-Hood and P1).  Elements are ordered along
// a Morton (Z-order) curve of their centroids and cut into patches of
// consecutive elements, so a patch is a compact piece of the mesh; inside a
// patch elements are in ascending id.  A patch's local DOFs (ascending global
// id) each own a run of slots -- one per element of the patch holding the
// DOF, in element order -- and elem_slot gives every (element, local k) its
// slot.  A DOF held by elements of one patch only is finished by that patch;
// any other DOF (on a seam between patches, or in no element) is summed from
// its per-patch partials, in patch order, by the seam pass.

struct fx_patches
{
  Index ndof = 0, num_elems = 0, num_parts = 0, max_patch_ldofs = 0, stride = 16;
  std::vector<Index>         patch_off, patch_elem, elem_vert, ldof_off, ldof_dof, ldof_part, seam_off,
      seam_dof;
  std::vector<std::uint16_t> elem_ldof, ldof_end;
};

namespace
{
// Spread the low 16 bits of v to the even bit positions.
std::uint32_t spread16(std::uint32_t v)
{
  v &= 0xffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}
} // namespace

extern "C" {

int fx_patches_create(Index ndof, Index num_elems, Index dofs_per_elem, const Index* dofs,
                      const Index* vconn3, Index num_vertices, const double* xv, const double* yv,
                      Index max_patch, fx_patches** out)
{
  return fx::guard([&] {
    fx::require(out && ndof >= 0 && num_elems >= 0 && num_vertices >= 0 && dofs_per_elem >= 1
                    && dofs_per_elem <= 16 && (num_elems == 0 || (dofs && vconn3 && xv && yv)),
                FX_ERR_ARG, "fx_patches_create: invalid arguments");
    const int per = static_cast<int>(dofs_per_elem), stride = (per + 3) / 4 * 4;
    fx::require(max_patch >= 1 && max_patch <= FX_TH_PATCH_MAX, FX_ERR_ARG,
                "fx_patches_create: max_patch must be in [1, FX_TH_PATCH_MAX]");
    *out        = nullptr;
    auto        p = std::make_unique<fx_patches>();
    const Index E = num_elems;
    p->ndof       = ndof;
    p->stride     = stride;
    p->num_elems  = E;
    for (Index e = 0; e < E; ++e)
    {
      for (int k = 0; k < per; ++k)
        fx::require(dofs[per * e + k] >= 0 && dofs[per * e + k] < ndof, FX_ERR_ARG,
                    "fx_patches_create: DOF id out of range");
      for (int k = 0; k < 3; ++k)
        fx::require(vconn3[3 * e + k] >= 0 && vconn3[3 * e + k] < num_vertices, FX_ERR_ARG,
                    "fx_patches_create: vertex id out of range");
    }
    // Morton order of the element centroids (ties by element id)
    double lo[2] = {0.0, 0.0}, hi[2] = {0.0, 0.0};
    std::vector<double> cx(E), cy(E);
    for (Index e = 0; e < E; ++e)
    {
      const Index* v = vconn3 + 3 * e;
      cx[e]          = (xv[v[0]] + xv[v[1]] + xv[v[2]]) / 3.0;
      cy[e]          = (yv[v[0]] + yv[v[1]] + yv[v[2]]) / 3.0;
      if (e == 0 || cx[e] < lo[0]) lo[0] = cx[e];
      if (e == 0 || cy[e] < lo[1]) lo[1] = cy[e];
      if (e == 0 || cx[e] > hi[0]) hi[0] = cx[e];
      if (e == 0 || cy[e] > hi[1]) hi[1] = cy[e];
    }
    const double span = std::max(hi[0] - lo[0], hi[1] - lo[1]);
    const double sc   = span > 0.0 ? 65535.0 / span : 0.0;
    std::vector<std::uint64_t> key(E);
    for (Index e = 0; e < E; ++e)
    {
      const auto qx = static_cast<std::uint32_t>(std::min(65535.0, std::max(0.0, (cx[e] - lo[0]) * sc)));
      const auto qy = static_cast<std::uint32_t>(std::min(65535.0, std::max(0.0, (cy[e] - lo[1]) * sc)));
      key[e] = (std::uint64_t(spread16(qx) | (spread16(qy) << 1)) << 32) | std::uint64_t(e);
    }
    std::sort(key.begin(), key.end());
    // Cut the curve into patches: a patch closes at max_patch elements, or
    // before an element that would give one of its DOFs a 17th element in
    // the patch (a DOF's rank within its run has 4 bits).
    std::vector<Index> stamp(ndof, -1), cnt(ndof, 0);
    p->patch_off.push_back(0);
    p->patch_elem.resize(E);
    for (Index i = 0; i < E; ++i)
    {
      const Index e  = static_cast<Index>(key[i] & 0xffffffffu);
      const Index q  = static_cast<Index>(p->patch_off.size()) - 1;
      bool        cut = i - p->patch_off.back() == max_patch;
      for (int k = 0; k < per && !cut; ++k)
      {
        const Index d = dofs[per * e + k];
        cut           = stamp[d] == q && cnt[d] == FX_TH_PATCH_RUN_MAX;
      }
      if (cut) p->patch_off.push_back(i);
      const Index qq = static_cast<Index>(p->patch_off.size()) - 1;
      for (int k = 0; k < per; ++k)
      {
        const Index d = dofs[per * e + k];
        if (stamp[d] != qq) stamp[d] = qq, cnt[d] = 0;
        ++cnt[d];
      }
      p->patch_elem[i] = e;
    }
    if (E > 0) p->patch_off.push_back(E);
    const Index P = static_cast<Index>(p->patch_off.size()) - 1;
    for (Index q = 0; q < P; ++q)
      std::sort(p->patch_elem.begin() + p->patch_off[q], p->patch_elem.begin() + p->patch_off[q + 1]);
    p->elem_vert.resize(3 * std::size_t(E));
    for (Index i = 0; i < E; ++i)
      for (int k = 0; k < 3; ++k) p->elem_vert[3 * std::size_t(i) + k] = vconn3[3 * p->patch_elem[i] + k];

    // patches touching each DOF, in patch order
    std::vector<Index> npatch(ndof, 0), last(ndof, -1);
    for (Index q = 0; q < P; ++q)
      for (Index i = p->patch_off[q]; i < p->patch_off[q + 1]; ++i)
        for (int k = 0; k < per; ++k)
        {
          const Index d = dofs[per * p->patch_elem[i] + k];
          if (last[d] != q) last[d] = q, ++npatch[d];
        }
    // seam DOFs: touched by several patches or by none; partials in patch order
    std::vector<Index> part_next(ndof, -1);
    p->seam_off.push_back(0);
    for (Index d = 0; d < ndof; ++d)
      if (npatch[d] != 1)
      {
        part_next[d] = p->seam_off.back();
        p->seam_dof.push_back(d);
        p->seam_off.push_back(p->seam_off.back() + npatch[d]);
      }
    p->num_parts = p->seam_off.back();
    fx::require(p->num_parts < FX_TH_PART_FIRST, FX_ERR_OVERFLOW,
                "fx_patches_create: too many seam partials");

    // local DOFs and runs of each patch
    p->elem_ldof.assign(stride * std::size_t(E), 0);
    p->ldof_off.push_back(0);
    std::vector<Index> local(ndof, -1), count, cursor, pdofs;
    for (Index q = 0; q < P; ++q)
    {
      const Index i0 = p->patch_off[q], i1 = p->patch_off[q + 1];
      pdofs.clear();
      for (Index i = i0; i < i1; ++i)
        for (int k = 0; k < per; ++k)
        {
          const Index d = dofs[per * p->patch_elem[i] + k];
          if (local[d] < 0) local[d] = 0, pdofs.push_back(d);
        }
      std::sort(pdofs.begin(), pdofs.end());
      const Index nl = static_cast<Index>(pdofs.size());
      fx::require(nl <= 4095 && nl <= per * FX_TH_PATCH_MAX, FX_ERR_UNSUPPORTED,
                  "fx_patches_create: patch has too many local DOFs");
      for (Index j = 0; j < nl; ++j) local[pdofs[j]] = j;
      count.assign(nl, 0);
      for (Index i = i0; i < i1; ++i)
        for (int k = 0; k < per; ++k)
        {
          const Index d = dofs[per * p->patch_elem[i] + k];
          for (int kk = 0; kk < k; ++kk)
            fx::require(dofs[per * p->patch_elem[i] + kk] != d, FX_ERR_UNSUPPORTED,
                        "fx_patches_create: an element lists one DOF twice");
          ++count[local[d]];
        }
      Index run = 0;
      for (Index j = 0; j < nl; ++j)
      {
        run += count[j];
        p->ldof_end.push_back(static_cast<std::uint16_t>(run));
        const Index d = pdofs[j];
        p->ldof_dof.push_back(d);
        if (npatch[d] == 1)
          p->ldof_part.push_back(-1);
        else
        {
          const bool first = part_next[d] == p->seam_off[std::lower_bound(p->seam_dof.begin(),
                                                                          p->seam_dof.end(), d)
                                                         - p->seam_dof.begin()];
          p->ldof_part.push_back(part_next[d]++ | (first ? FX_TH_PART_FIRST : 0));
        }
      }
      cursor.assign(nl, 0); // rank of the next element in each run
      for (Index i = i0; i < i1; ++i)
        for (int k = 0; k < per; ++k)
        {
          const Index j = local[dofs[per * p->patch_elem[i] + k]];
          p->elem_ldof[stride * std::size_t(i) + k] =
              static_cast<std::uint16_t>(j | (cursor[j]++ << 12));
        }
      for (Index d : pdofs) local[d] = -1;
      p->ldof_off.push_back(p->ldof_off.back() + nl);
      p->max_patch_ldofs = std::max(p->max_patch_ldofs, nl);
    }
    *out = p.release();
  });
}

int fx_patches_info_get(const fx_patches* p, fx_patches_info* info)
{
  return fx::guard([&] {
    fx::require(p && info, FX_ERR_ARG, "fx_patches_info_get: null argument");
    info->num_patches = static_cast<Index>(p->patch_off.size()) - 1;
    info->num_ldofs   = static_cast<Index>(p->ldof_dof.size());
    info->num_seam    = static_cast<Index>(p->seam_dof.size());
    info->num_parts   = p->num_parts;
    info->max_patch_ldofs = p->max_patch_ldofs;
    info->ldof_stride     = p->stride;
  });
}

int fx_patches_export(const fx_patches* p, Index* patch_off, Index* patch_elem,
                         Index* elem_vert, unsigned short* elem_ldof, Index* ldof_off,
                         Index* ldof_dof,
                         Index* ldof_part, unsigned short* ldof_end, Index* seam_off,
                         Index* seam_dof)
{
  return fx::guard([&] {
    fx::require(p && patch_off && ldof_off && seam_off
                    && (p->num_elems == 0 || (patch_elem && elem_vert && elem_ldof))
                    && (p->ldof_dof.empty() || (ldof_dof && ldof_part && ldof_end))
                    && (p->seam_dof.empty() || seam_dof),
                FX_ERR_ARG, "fx_patches_export: null array");
    std::copy(p->patch_off.begin(), p->patch_off.end(), patch_off);
    std::copy(p->patch_elem.begin(), p->patch_elem.end(), patch_elem);
    std::copy(p->elem_vert.begin(), p->elem_vert.end(), elem_vert);
    std::copy(p->elem_ldof.begin(), p->elem_ldof.end(), elem_ldof);
    std::copy(p->ldof_off.begin(), p->ldof_off.end(), ldof_off);
    std::copy(p->ldof_dof.begin(), p->ldof_dof.end(), ldof_dof);
    std::copy(p->ldof_part.begin(), p->ldof_part.end(), ldof_part);
    std::copy(p->ldof_end.begin(), p->ldof_end.end(), ldof_end);
    std::copy(p->seam_off.begin(), p->seam_off.end(), seam_off);
    std::copy(p->seam_dof.begin(), p->seam_dof.end(), seam_dof);
  });
}

void fx_patches_destroy(fx_patches* p) { delete p; }

} // extern "C"

// P1 assembly plan: patches of at most max_patch elements in which every row
// is finished by the one patch that owns its node.  Elements are visited in
// Morton order of their centroids; a node is owned by the patch that meets it
// first, and that patch takes all of the node's elements (its own and, as a
// halo, its neighbours').  A node is ownable when its row has at most 16
// entries and it has 1 to 16 elements; the others (and nodes in no element)
// are left to the row-per-lane kernel.
struct fx_p1_asm
{
  Index num_nodes = 0, max_patch_ldofs = 0, max_tile = 0, max_runs = 0;
  std::vector<Index>         patch_off{0}, patch_elem, ldof_off{0}, ldof_dof, patch_tile, run_off{0},
      runs, rows_left;
  std::vector<std::uint16_t> elem_ldof, elem_pos, ldof_tile;
};

extern "C" {

int fx_p1_asm_create(Index num_nodes, Index num_elems, const Index* conn, const double* x,
                     const double* y, const Index* row_ptr, const Index* col_ind, Index max_patch,
                     Index max_tile, fx_p1_asm** out)
{
  return fx::guard([&] {
    fx::require(out && num_nodes >= 0 && num_elems >= 0 && row_ptr
                    && (num_elems == 0 || (conn && x && y && col_ind)),
                FX_ERR_ARG, "fx_p1_asm_create: invalid arguments");
    fx::require(max_patch >= 1 && max_patch <= FX_TH_PATCH_MAX && max_tile >= 0
                    && max_tile < 0xffff,
                FX_ERR_ARG, "fx_p1_asm_create: max_patch must be in [1, FX_TH_PATCH_MAX], max_tile below 65535");
    *out         = nullptr;
    auto p       = std::make_unique<fx_p1_asm>();
    const Index N = num_nodes, E = num_elems;
    p->num_nodes = N;
    for (Index i = 0; i < 3 * E; ++i)
      fx::require(conn[i] >= 0 && conn[i] < N, FX_ERR_ARG, "fx_p1_asm_create: node id out of range");
    // each node's elements, ascending (counting sort)
    std::vector<Index> noff(N + 1, 0), nel(3 * std::size_t(E));
    for (Index i = 0; i < 3 * E; ++i) ++noff[conn[i] + 1];
    for (Index n = 0; n < N; ++n) noff[n + 1] += noff[n];
    {
      std::vector<Index> cur(noff.begin(), noff.end() - 1);
      for (Index e = 0; e < E; ++e)
        for (int k = 0; k < 3; ++k) nel[cur[conn[3 * e + k]]++] = e;
    }
    auto rank_of = [&](Index n, Index e) {
      return static_cast<Index>(std::lower_bound(nel.begin() + noff[n], nel.begin() + noff[n + 1], e)
                                - (nel.begin() + noff[n]));
    };
    auto rowlen = [&](Index n) { return row_ptr[n + 1] - row_ptr[n]; };
    std::vector<char> ownable(N);
    for (Index n = 0; n < N; ++n)
    {
      const Index ne = noff[n + 1] - noff[n];
      ownable[n]     = ne >= 1 && ne <= 16 && rowlen(n) <= 16 && rowlen(n) + 0 <= max_tile;
    }
    // Morton order of the element centroids (ties by element id)
    std::vector<std::uint64_t> key(E);
    double lo[2] = {0.0, 0.0}, hi[2] = {0.0, 0.0};
    std::vector<double> cx(E), cy(E);
    for (Index e = 0; e < E; ++e)
    {
      cx[e] = (x[conn[3 * e]] + x[conn[3 * e + 1]] + x[conn[3 * e + 2]]) / 3.0;
      cy[e] = (y[conn[3 * e]] + y[conn[3 * e + 1]] + y[conn[3 * e + 2]]) / 3.0;
      if (e == 0 || cx[e] < lo[0]) lo[0] = cx[e];
      if (e == 0 || cy[e] < lo[1]) lo[1] = cy[e];
      if (e == 0 || cx[e] > hi[0]) hi[0] = cx[e];
      if (e == 0 || cy[e] > hi[1]) hi[1] = cy[e];
    }
    const double span = std::max(hi[0] - lo[0], hi[1] - lo[1]);
    const double sc   = span > 0.0 ? 65535.0 / span : 0.0;
    for (Index e = 0; e < E; ++e)
    {
      const auto qx = static_cast<std::uint32_t>(std::min(65535.0, std::max(0.0, (cx[e] - lo[0]) * sc)));
      const auto qy = static_cast<std::uint32_t>(std::min(65535.0, std::max(0.0, (cy[e] - lo[1]) * sc)));
      key[e] = (std::uint64_t(spread16(qx) | (spread16(qy) << 1)) << 32) | std::uint64_t(e);
    }
    std::sort(key.begin(), key.end());

    std::vector<Index> owner(N, -1), estamp(E, -1), local(N, -1);
    std::vector<Index> ext, owned, nodes, fresh;
    Index              tile = 0, q = 0;
    auto close = [&]() {
      if (owned.empty()) return;
      std::sort(ext.begin(), ext.end());
      nodes.clear();
      for (Index e : ext)
        for (int k = 0; k < 3; ++k)
        {
          const Index n = conn[3 * e + k];
          if (local[n] < 0) local[n] = 0, nodes.push_back(n);
        }
      std::sort(nodes.begin(), nodes.end());
      const Index nl = static_cast<Index>(nodes.size());
      for (Index j = 0; j < nl; ++j) local[nodes[j]] = j;
      // owned rows in ascending node order: tile offsets and runs of consecutive ids
      Index t = 0, prev = -2, r0 = static_cast<Index>(p->runs.size() / 2);
      for (Index n : nodes)
      {
        std::uint16_t tf = 0xffff;
        if (owner[n] == q)
        {
          if (n != prev + 1) p->runs.push_back(t), p->runs.push_back(row_ptr[n]);
          tf   = static_cast<std::uint16_t>(t);
          t   += rowlen(n);
          prev = n;
        }
        else
          prev = -2;
        p->ldof_tile.push_back(tf);
        p->ldof_dof.push_back(n);
      }
      int steps = 0;
      for (Index e : ext)
      {
        p->patch_elem.push_back(e);
        for (int k = 0; k < 4; ++k)
        {
          std::uint16_t w = 0, pw = 0;
          if (k < 3)
          {
            const Index n = conn[3 * e + k];
            w             = static_cast<std::uint16_t>(local[n]);
            if (owner[n] == q)
            {
              const Index rk = rank_of(n, e);
              steps          = std::max(steps, static_cast<int>(rk) + 1);
              w              = static_cast<std::uint16_t>(w | (rk << 12));
              const Index* b = col_ind + row_ptr[n];
              const Index* f = col_ind + row_ptr[n + 1];
              for (int c = 0; c < 3; ++c)
              {
                const Index* at = std::lower_bound(b, f, conn[3 * e + c]);
                fx::require(at != f && *at == conn[3 * e + c], FX_ERR_LAYOUT,
                            "fx_p1_asm_create: an element column is missing from its row");
                pw = static_cast<std::uint16_t>(pw | ((at - b) << (4 * c)));
              }
            }
          }
          p->elem_ldof.push_back(w);
          p->elem_pos.push_back(pw);
        }
      }
      for (Index n : nodes) local[n] = -1;
      p->patch_off.push_back(static_cast<Index>(p->patch_elem.size()));
      p->ldof_off.push_back(static_cast<Index>(p->ldof_dof.size()));
      p->patch_tile.push_back(t | (steps << 20));
      p->run_off.push_back(static_cast<Index>(p->runs.size() / 2));
      p->max_patch_ldofs = std::max(p->max_patch_ldofs, nl);
      p->max_tile        = std::max(p->max_tile, t);
      p->max_runs        = std::max(p->max_runs, p->run_off.back() - r0);
      ext.clear();
      owned.clear();
      tile = 0;
      ++q;
    };
    for (Index i = 0; i < E; ++i)
    {
      const Index e = static_cast<Index>(key[i] & 0xffffffffu);
      for (int attempt = 0; attempt < 2; ++attempt)
      {
        // nodes of e nobody owns yet, and the elements they would bring
        fresh.clear();
        Index add = 0, rows = 0;
        for (int k = 0; k < 3; ++k)
        {
          const Index n = conn[3 * e + k]; // an element's three nodes are distinct
          if (owner[n] >= 0 || !ownable[n]) continue;
          fresh.push_back(n);
          rows += rowlen(n);
          for (Index t = noff[n]; t < noff[n + 1]; ++t)
            if (estamp[nel[t]] != q && estamp[nel[t]] != -2 - i) estamp[nel[t]] = -2 - i, ++add;
        }
        // undo the trial marks
        for (Index n : fresh)
          for (Index t = noff[n]; t < noff[n + 1]; ++t)
            if (estamp[nel[t]] == -2 - i) estamp[nel[t]] = -1;
        if (fresh.empty()) break;
        const bool fits = static_cast<Index>(ext.size()) + add <= max_patch && tile + rows <= max_tile;
        if (!fits && !owned.empty() && attempt == 0)
        {
          close(); // and try again in a new, empty patch
          continue;
        }
        if (!fits)
        {
          // too big for an empty patch: these nodes are left to the row kernel
          for (Index n : fresh) ownable[n] = 0;
          break;
        }
        for (Index n : fresh)
        {
          owner[n] = q;
          owned.push_back(n);
          for (Index t = noff[n]; t < noff[n + 1]; ++t)
            if (estamp[nel[t]] != q) estamp[nel[t]] = q, ext.push_back(nel[t]);
        }
        tile += rows;
        break;
      }
    }
    close();
    for (Index n = 0; n < N; ++n)
      if (owner[n] < 0) p->rows_left.push_back(n);
    *out = p.release();
  });
}

int fx_p1_asm_info_get(const fx_p1_asm* p, fx_p1_asm_info* info)
{
  return fx::guard([&] {
    fx::require(p && info, FX_ERR_ARG, "fx_p1_asm_info_get: null argument");
    info->num_patches     = static_cast<Index>(p->patch_off.size()) - 1;
    info->num_entries     = static_cast<Index>(p->patch_elem.size());
    info->num_ldofs       = static_cast<Index>(p->ldof_dof.size());
    info->num_runs        = static_cast<Index>(p->runs.size() / 2);
    info->num_left        = static_cast<Index>(p->rows_left.size());
    info->max_patch_ldofs = p->max_patch_ldofs;
    info->max_tile        = p->max_tile;
    info->max_runs        = p->max_runs;
  });
}

int fx_p1_asm_export(const fx_p1_asm* p, Index* patch_off, Index* patch_elem,
                     unsigned short* elem_ldof, unsigned short* elem_pos, Index* ldof_off,
                     Index* ldof_dof, unsigned short* ldof_tile, Index* patch_tile, Index* run_off,
                     Index* runs, Index* rows_left)
{
  return fx::guard([&] {
    fx::require(p && patch_off && ldof_off && run_off
                    && (p->patch_elem.empty() || (patch_elem && elem_ldof && elem_pos))
                    && (p->ldof_dof.empty() || (ldof_dof && ldof_tile))
                    && (p->patch_tile.empty() || patch_tile) && (p->runs.empty() || runs)
                    && (p->rows_left.empty() || rows_left),
                FX_ERR_ARG, "fx_p1_asm_export: null array");
    std::copy(p->patch_off.begin(), p->patch_off.end(), patch_off);
    std::copy(p->patch_elem.begin(), p->patch_elem.end(), patch_elem);
    std::copy(p->elem_ldof.begin(), p->elem_ldof.end(), elem_ldof);
    std::copy(p->elem_pos.begin(), p->elem_pos.end(), elem_pos);
    std::copy(p->ldof_off.begin(), p->ldof_off.end(), ldof_off);
    std::copy(p->ldof_dof.begin(), p->ldof_dof.end(), ldof_dof);
    std::copy(p->ldof_tile.begin(), p->ldof_tile.end(), ldof_tile);
    std::copy(p->patch_tile.begin(), p->patch_tile.end(), patch_tile);
    std::copy(p->run_off.begin(), p->run_off.end(), run_off);
    std::copy(p->runs.begin(), p->runs.end(), runs);
    std::copy(p->rows_left.begin(), p->rows_left.end(), rows_left);
  });
}

void fx_p1_asm_destroy(fx_p1_asm* p) { delete p; }

// ---------------------------------------------------------------------------
// CSR transpose permutation: counting sort by column, rows visited in order.

int fx_csr_transpose_map(Index rows, Index cols, const Index* row_ptr, const Index* col_ind,
                         Index* t_ptr, Index* t_row, Index* t_perm)
{
  return fx::guard([&] {
    fx::require(rows >= 0 && cols >= 0 && row_ptr && t_ptr, FX_ERR_ARG,
                "fx_csr_transpose_map: invalid arguments");
    const Index nnz = row_ptr[rows];
    fx::require(nnz == 0 || (col_ind && t_row && t_perm), FX_ERR_ARG,
                "fx_csr_transpose_map: null arrays");
    std::fill(t_ptr, t_ptr + cols + 1, 0);
    for (Index k = 0; k < nnz; ++k)
    {
      fx::require(col_ind[k] >= 0 && col_ind[k] < cols, FX_ERR_ARG,
                  "fx_csr_transpose_map: column index out of range");
      ++t_ptr[col_ind[k] + 1];
    }
    for (Index c = 0; c < cols; ++c) t_ptr[c + 1] += t_ptr[c];
    std::vector<Index> next(t_ptr, t_ptr + cols);
    for (Index r = 0; r < rows; ++r)
      for (Index k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
      {
        const Index d = next[col_ind[k]]++;
        t_row[d]      = r;
        t_perm[d]     = k;
      }
  });
}

// ---------------------------------------------------------------------------
// Stencil blocks of the multi-vector product: a full block of 64 rows whose
// rows all have the same count S (1..max_len) of entries at the same column
// offsets (col - row, entry by entry) gets the code (start << 5) | S, start
// indexing its offsets in st_off (equal stencils share one copy); any other
// block gets -1.

int fx_csr_stencil_plan(Index rows, Index cols, const Index* row_ptr, const Index* col_ind,
                        Index max_len, Index* num_blocks, Index* num_offsets, Index* blk_code,
                        Index* st_off)
{
  return fx::guard([&] {
    fx::require(rows >= 0 && cols >= 0 && row_ptr && num_blocks && num_offsets && max_len >= 1
                    && max_len <= 31,
                FX_ERR_ARG, "fx_csr_stencil_plan: invalid arguments");
    fx::require(row_ptr[rows] == 0 || col_ind, FX_ERR_ARG, "fx_csr_stencil_plan: null column array");
    const Index                     nb = (rows + 63) / 64;
    std::map<std::vector<Index>, Index> seen;
    std::vector<Index>              flat, off;
    for (Index b = 0; b < nb; ++b)
    {
      const Index r0 = 64 * b;
      Index       code = -1;
      const Index S    = row_ptr[r0 + 1] - row_ptr[r0]; // r0 < rows
      if (r0 + 64 <= rows && S >= 1 && S <= max_len)
      {
        off.resize(static_cast<std::size_t>(S));
        for (Index k = 0; k < S; ++k) off[k] = col_ind[row_ptr[r0] + k] - r0;
        bool same = true;
        for (Index r = r0 + 1; r < r0 + 64 && same; ++r)
        {
          same = row_ptr[r + 1] - row_ptr[r] == S;
          for (Index k = 0; k < S && same; ++k) same = col_ind[row_ptr[r] + k] - r == off[k];
        }
        if (same)
        {
          auto it = seen.find(off);
          if (it == seen.end())
          {
            fx::require(flat.size() + off.size() < (std::size_t(1) << 26), FX_ERR_UNSUPPORTED,
                        "fx_csr_stencil_plan: too many distinct stencils");
            it = seen.emplace(off, static_cast<Index>(flat.size())).first;
            flat.insert(flat.end(), off.begin(), off.end());
          }
          code = (it->second << 5) | S;
        }
      }
      if (blk_code) blk_code[b] = code;
    }
    if (st_off) std::copy(flat.begin(), flat.end(), st_off);
    *num_blocks  = nb;
    *num_offsets = static_cast<Index>(flat.size());
  });
}

// ---------------------------------------------------------------------------
// Stencil blocks of the P1 assembly.  A full block of 64 rows is uniform when
// its rows have the same column offsets (S <= max_len entries), the same
// number T <= max_len of incidences with the same local row and column
// positions, incidence by incidence, and elements e_k(r) = A r + c_k for one
// A and per-incidence constants c_k, and when its bands (runs of
// consecutive offsets; a band of w offsets spans 63 + w values of a vector)
// fit max_band values in all.  Its record, shared by equal blocks:
// [S, T, A, offset_0 .. offset_(S-1), sidx_0 .. sidx_(S-1), (c_k, meta_k)
// for k < T], meta as in the row incidence (li | p0 << 8 | p1 << 16 | p2 <<
// 24) and sidx_q the index of slot q's lane-0 value in the block's band
// buffer (bands packed in order); blk_code is the record's first word, or -1.

int fx_p1_stencil_plan(Index num_nodes, const Index* row_ptr, const Index* col_ind,
                       const Index* inc_off, const Index* inc, Index max_len, Index max_band,
                       Index* num_blocks, Index* num_words, Index* blk_code, Index* words)
{
  return fx::guard([&] {
    fx::require(num_nodes >= 0 && row_ptr && inc_off && num_blocks && num_words && max_len >= 1
                    && max_len <= 255 && max_band >= 64,
                FX_ERR_ARG, "fx_p1_stencil_plan: invalid arguments");
    fx::require((row_ptr[num_nodes] == 0 || col_ind) && (inc_off[num_nodes] == 0 || inc),
                FX_ERR_ARG, "fx_p1_stencil_plan: null array");
    const Index                         nb = (num_nodes + 63) / 64;
    std::map<std::vector<Index>, Index> seen;
    std::vector<Index>                  flat, rec;
    for (Index b = 0; b < nb; ++b)
    {
      const Index r0 = 64 * b;
      Index       code = -1;
      const Index S = row_ptr[r0 + 1] - row_ptr[r0], T = inc_off[r0 + 1] - inc_off[r0];
      bool same = r0 + 64 <= num_nodes && S >= 1 && S <= max_len && T >= 1 && T <= max_len
                  && inc_off[r0 + 2] - inc_off[r0 + 1] == T;
      std::int64_t A = 0;
      if (same)
      {
        const auto elem = [&](Index r, Index k) { return std::int64_t(inc[2 * (inc_off[r] + k)]); };
        const auto meta = [&](Index r, Index k) { return inc[2 * (inc_off[r] + k) + 1]; };
        A = elem(r0 + 1, 0) - elem(r0, 0);
        rec.assign({S, T, 0});
        for (Index k = 0; k < S; ++k) rec.push_back(col_ind[row_ptr[r0] + k] - r0);
        // band buffer: a band of w consecutive offsets takes 63 + w values
        Index used = 0;
        for (Index q = 0; q < S;)
        {
          Index w = 1;
          while (q + w < S && rec[3 + q + w] == rec[3 + q] + w) ++w;
          for (Index d = 0; d < w; ++d) rec.push_back(used + d);
          used += 63 + w;
          q += w;
        }
        same = used <= max_band;
        for (Index k = 0; k < T; ++k)
        {
          const std::int64_t c = elem(r0, k) - A * r0;
          same = same && c >= std::numeric_limits<Index>::min()
                 && c <= std::numeric_limits<Index>::max();
          rec.push_back(static_cast<Index>(c));
          rec.push_back(meta(r0, k));
        }
        same = same && A >= std::numeric_limits<Index>::min() && A <= std::numeric_limits<Index>::max();
        for (Index r = r0 + 1; r < r0 + 64 && same; ++r)
        {
          same = row_ptr[r + 1] - row_ptr[r] == S && inc_off[r + 1] - inc_off[r] == T;
          for (Index k = 0; k < S && same; ++k) same = col_ind[row_ptr[r] + k] - r == rec[3 + k];
          for (Index k = 0; k < T && same; ++k)
            same = elem(r, k) - A * r == rec[3 + 2 * S + 2 * k]
                   && meta(r, k) == rec[4 + 2 * S + 2 * k];
        }
      }
      if (same)
      {
        rec[2]  = static_cast<Index>(A);
        auto it = seen.find(rec);
        if (it == seen.end())
        {
          fx::require(flat.size() + rec.size() <= std::size_t(std::numeric_limits<Index>::max()),
                      FX_ERR_UNSUPPORTED, "fx_p1_stencil_plan: too many distinct stencils");
          it = seen.emplace(rec, static_cast<Index>(flat.size())).first;
          flat.insert(flat.end(), rec.begin(), rec.end());
        }
        code = it->second;
      }
      if (blk_code) blk_code[b] = code;
    }
    if (words) std::copy(flat.begin(), flat.end(), words);
    *num_blocks = nb;
    *num_words  = static_cast<Index>(flat.size());
  });
}

// ---------------------------------------------------------------------------
// Boundary map.  Validation order and messages follow femx's makeBoundaryMap
// (femx/assembly/BoundaryMap.cpp:59-139).

int fx_bcmap_create(Index rows, const Index* row_ptr, const Index* col_ind, Index num_bcs,
                    const Index* dofs, fx_bcmap** out)
{
  return fx::guard([&] {
    fx::require(out != nullptr, FX_ERR_ARG, "fx_bcmap_create: null output handle");
    *out = nullptr;
    fx::require(rows >= 0 && num_bcs >= 0 && row_ptr && (num_bcs == 0 || dofs), FX_ERR_ARG,
                "fx_bcmap_create: invalid arguments");
    auto b  = std::make_unique<fx_bcmap>();
    b->rows = rows;
    b->bc_rows.resize(num_bcs);
    b->diag.assign(num_bcs, -1);
    b->col_off.assign(num_bcs + 1, 0);
    b->mask.assign(rows, 0);
    std::vector<Index> bc_by_col(rows, -1);
    for (Index ib = 0; ib < num_bcs; ++ib)
    {
      const Index dof = dofs[ib];
      fx::require(dof >= 0 && dof < rows, FX_ERR_ARG, "BoundaryMap constrained DOF is out of range");
      fx::require(b->mask[dof] == 0, FX_ERR_ARG, "BoundaryMap constrained DOFs must be unique");
      b->bc_rows[ib] = dof;
      b->mask[dof]   = 1;
      bc_by_col[dof] = ib;
    }
    for (Index r = 0; r < rows; ++r)
      for (Index k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
      {
        const Index c = col_ind[k];
        fx::require(c >= 0 && c < rows, FX_ERR_ARG, "BoundaryMap requires a square CSR pattern");
        const Index ib = bc_by_col[c];
        if (ib >= 0)
        {
          ++b->col_off[ib + 1];
          if (r == c)
          {
            fx::require(b->diag[ib] < 0, FX_ERR_PATTERN,
                        "BoundaryMap constrained row has duplicate diagonal entries");
            b->diag[ib] = k;
          }
        }
      }
    for (Index ib = 0; ib < num_bcs; ++ib)
    {
      fx::require(b->diag[ib] >= 0, FX_ERR_PATTERN, "BoundaryMap constrained row has no diagonal entry");
      b->col_off[ib + 1] += b->col_off[ib];
    }
    const Index nce = b->col_off[num_bcs];
    b->col_ent.resize(nce);
    b->col_rows.resize(nce);
    std::vector<Index> next(b->col_off.begin(), b->col_off.end() - 1);
    for (Index r = 0; r < rows; ++r)
      for (Index k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
      {
        const Index ib = bc_by_col[col_ind[k]];
        if (ib >= 0)
        {
          const Index d = next[ib]++;
          b->col_ent[d]  = k;
          b->col_rows[d] = r;
        }
      }
    // Row-grouped elimination lists: entries of unconstrained rows in
    // constrained columns, each row's entries in ascending boundary index.
    std::vector<Index> cnt(rows + 1, 0);
    for (Index i = 0; i < nce; ++i)
      if (!b->mask[b->col_rows[i]]) ++cnt[b->col_rows[i] + 1];
    for (Index r = 0; r < rows; ++r) cnt[r + 1] += cnt[r];
    const Index ne = cnt[rows];
    b->elim_k.resize(ne);
    b->elim_ib.resize(ne);
    std::vector<Index> pos(cnt.begin(), cnt.end() - 1);
    for (Index ib = 0; ib < num_bcs; ++ib)
      for (Index i = b->col_off[ib]; i < b->col_off[ib + 1]; ++i)
      {
        const Index r = b->col_rows[i];
        if (b->mask[r]) continue;
        const Index d = pos[r]++;
        b->elim_k[d]  = b->col_ent[i];
        b->elim_ib[d] = ib;
      }
    b->elim_off.push_back(0);
    for (Index r = 0; r < rows; ++r)
      if (cnt[r + 1] > cnt[r])
      {
        b->elim_rows.push_back(r);
        b->elim_off.push_back(cnt[r + 1]);
      }
    *out = b.release();
  });
}

int fx_bcmap_info_get(const fx_bcmap* b, fx_bcmap_info* info)
{
  return fx::guard([&] {
    fx::require(b && info, FX_ERR_ARG, "fx_bcmap_info_get: null argument");
    info->rows             = b->rows;
    info->num_bcs          = static_cast<Index>(b->bc_rows.size());
    info->num_col_entries  = static_cast<Index>(b->col_ent.size());
    info->num_elim_rows    = static_cast<Index>(b->elim_rows.size());
    info->num_elim_entries = static_cast<Index>(b->elim_k.size());
  });
}

int fx_bcmap_export(const fx_bcmap* b, Index* bc_rows, Index* diag, Index* col_offsets,
                    Index* col_entries, Index* col_rows, Index* bc_mask)
{
  return fx::guard([&] {
    fx::require(b != nullptr, FX_ERR_ARG, "fx_bcmap_export: null map");
    if (bc_rows) std::copy(b->bc_rows.begin(), b->bc_rows.end(), bc_rows);
    if (diag) std::copy(b->diag.begin(), b->diag.end(), diag);
    if (col_offsets) std::copy(b->col_off.begin(), b->col_off.end(), col_offsets);
    if (col_entries) std::copy(b->col_ent.begin(), b->col_ent.end(), col_entries);
    if (col_rows) std::copy(b->col_rows.begin(), b->col_rows.end(), col_rows);
    if (bc_mask) std::copy(b->mask.begin(), b->mask.end(), bc_mask);
  });
}

int fx_bcmap_export_elim(const fx_bcmap* b, Index* elim_rows, Index* elim_off, Index* elim_k,
                         Index* elim_ib)
{
  return fx::guard([&] {
    fx::require(b && elim_off, FX_ERR_ARG, "fx_bcmap_export_elim: null argument");
    if (elim_rows) std::copy(b->elim_rows.begin(), b->elim_rows.end(), elim_rows);
    std::copy(b->elim_off.begin(), b->elim_off.end(), elim_off);
    if (elim_k) std::copy(b->elim_k.begin(), b->elim_k.end(), elim_k);
    if (elim_ib) std::copy(b->elim_ib.begin(), b->elim_ib.end(), elim_ib);
  });
}

void fx_bcmap_destroy(fx_bcmap* b) { delete b; }

} // extern "C"
