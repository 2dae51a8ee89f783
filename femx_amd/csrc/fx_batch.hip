// Batched assembly and adjoint VJP of many P1 Poisson and Taylor-Hood
// Navier-Stokes instances of uneven size.
//
// femx assembles one problem per call (femx/assembly/CudaAssembly.hpp:
// 319-364), each with its own resize, zero, attribute query and launch.  Here
// a batch owns device descriptors of all instances and one work list per
// phase -- (instance, block) pairs covering every row block of every
// instance.  Each phase is one persistent launch: a grid sized to the GPU
// (compute units x the larger of a tunable multiplier and the phase kernel's
// resident workgroups per compute unit) whose workgroups walk the work list
// -- each XCD's workgroups one contiguous eighth of it, so neighbouring
// blocks share an L2 -- running the same per-block device code as the
// single-instance kernels, so each instance's results are bitwise equal to
// its plain launch.  The phases form four independent chains -- P1 assembly,
// Taylor-Hood assembly, and the P1 and Taylor-Hood VJPs (each a patch then a
// seam pass) -- launched on private branch streams forked from the caller's
// stream, so they run concurrently; captured, they become the branches of
// one hipGraph, replayed with a single hipGraphLaunch.  Several batches (for
// instance one per size class, each with its own grid and row tile) can be
// captured into one graph together (fx_batch_graph_create).  With
// grid_per_cu = 0 a phase launches one workgroup per work item instead of a
// persistent grid, so that concurrent phases share the compute units as
// their workgroups retire.
#include <algorithm>
#include <memory>
#include <vector>

#include "fx_gls_dev.hpp"
#include "fx_p1_dev.hpp"
#include "fx_th_dev.hpp"

struct fx_batch
{
  int         num_p1 = 0, num_th = 0, num_gls = 0;
  int         grid   = 0;       // compute units x the tunable multiplier
  int         pgrid[9] = {};    // per phase: at least one full residency
  int         cus    = 0;
  std::size_t p1_lds = 0, th_lds = 0, th_vjp_lds = 0;
  int         th_max_ldofs = 0; // the most local DOFs of a VJP patch over the instances
  int         p1_max_ldofs = 0;
  std::size_t p1_vjp_lds   = 0;
  int         th_tile = 0; // row tile per wave of the Taylor-Hood assembly (doubles)
  bool        p1_long = false; // some P1 row is longer than the LDS tile width
  bool        th_rounds = false; // every Taylor-Hood instance has round tables
  std::size_t gls_lds = 0, gls_vjp_lds = 0;
  int         gls_max_ldofs = 0;
  // device memory
  fxk::P1AsmArgs*  d_p1_asm  = nullptr;
  fxk::P1VjpArgs*  d_p1_vjp  = nullptr;
  fxk::PatchArgs*  d_p1_pat  = nullptr;
  fxk::SeamArgs*   d_p1_seam = nullptr;
  fxk::ThAsmArgs*  d_th_asm  = nullptr;
  fxk::ThVjpArgs*   d_th_vjp  = nullptr;
  fxk::PatchArgs* d_th_pat  = nullptr;
  fxk::SeamArgs*  d_th_seam = nullptr;
  fxk::GlsAsmArgs*   d_gls_asm  = nullptr;
  fxk::GlsVjpArgs*   d_gls_vjp  = nullptr;
  fxk::PatchArgs*    d_gls_pat  = nullptr;
  fxk::SeamArgs*     d_gls_seam = nullptr;
  fxk::GlsPatchOut*  d_gls_out  = nullptr;
  int2*            d_work[9] = {};
  int              nwork[9]  = {};
  hipStream_t      cap_stream = nullptr;
  hipGraph_t       graph      = nullptr;
  hipGraphExec_t   exec       = nullptr;
  int              graph_what = 0;
  int              device     = 0;
  // independent chains (P1 assembly, TH assembly, P1 VJP, TH VJP, GLS
  // assembly, GLS VJP) run on private branch streams forked from the
  // caller's stream and joined back
  hipStream_t branch[6] = {};
  hipEvent_t  fork      = nullptr;
  hipEvent_t  done[6]   = {};

  ~fx_batch()
  {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    for (auto& st : branch)
      if (st) (void)hipStreamDestroy(st);
    if (fork) (void)hipEventDestroy(fork);
    for (auto& e : done)
      if (e) (void)hipEventDestroy(e);
    for (void* p : {static_cast<void*>(d_p1_asm), static_cast<void*>(d_p1_vjp),
                    static_cast<void*>(d_th_asm), static_cast<void*>(d_th_vjp),
                    static_cast<void*>(d_th_pat), static_cast<void*>(d_th_seam),
                    static_cast<void*>(d_p1_pat), static_cast<void*>(d_p1_seam),
                    static_cast<void*>(d_gls_asm), static_cast<void*>(d_gls_vjp),
                    static_cast<void*>(d_gls_pat), static_cast<void*>(d_gls_seam),
                    static_cast<void*>(d_gls_out)})
      if (p) (void)hipFree(p);
    for (auto* w : d_work)
      if (w) (void)hipFree(w);
  }
};

struct fx_batch_graph
{
  hipStream_t    origin = nullptr;
  hipGraph_t     graph  = nullptr;
  hipGraphExec_t exec   = nullptr;
  ~fx_batch_graph()
  {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (origin) (void)hipStreamDestroy(origin);
  }
};

namespace
{
using namespace fxk;

enum Phase
{
  kP1Asm = 0,
  kThAsm = 1,
  kP1Vjp = 2,
  kThVjpPatch = 3,
  kThVjpSeam = 4,
  kP1VjpSeam = 5,
  kGlsAsm = 6,
  kGlsVjpPatch = 7,
  kGlsVjpSeam = 8
};
constexpr int kPhases = 9, kChains = 6;
// blocks per Taylor-Hood assembly work item when an instance has round
// tables (256-instance assembly graph: 2: 5.82 ms, 4: 5.63, 6: 5.62, 8: 5.56,
// 12: 5.49, 16: 5.52; capping an instance's walks at 1/64 of its blocks, so
// that small instances spread over more waves, took 5.69)
constexpr int kThRtBatchWalk = 12;

// capped at 96 VGPRs (5 waves/SIMD, as k_p1_assemble reaches uncapped)
template <bool kLong>
__global__ __launch_bounds__(kP1Block) __attribute__((amdgpu_waves_per_eu(5))) void
k_batch_p1_assemble(const P1AsmArgs* __restrict__ inst, const int2* __restrict__ work, int nwork)
{
  extern __shared__ double smem[];
  const fx::XcdWalk wk = fx::xcd_walk(nwork);
  for (int w = wk.first; w < wk.end; w += wk.stride)
  {
    const int2 it = work[w];
    p1_assemble_block<kLong>(it.y, inst[it.x], smem);
  }
}

__global__ __launch_bounds__(kPatchLanes) void k_batch_p1_vjp_patch(const P1VjpArgs* __restrict__ inst,
                                                                    const PatchArgs* __restrict__ pat,
                                                                    const int2* __restrict__ work,
                                                                    int nwork, int max_patch_ldofs)
{
  extern __shared__ __align__(16) unsigned char lds[];
  const fx::XcdWalk wk = fx::xcd_walk(nwork);
  for (int w = wk.first; w < wk.end; w += wk.stride)
  {
    const int2 it = work[w];
    p1_vjp_patch_block(it.y, inst[it.x], pat[it.x], lds, max_patch_ldofs);
  }
}

// Each wave walks the work list on its own (xcd_walk with one walker per
// wave) with its own row tile of `tile` doubles; the workgroup shares the
// tables.
// At most 256 VGPRs (2 waves per SIMD): a register count past 256 would
// halve the occupancy, which costs far more than a few spills.
// Rounds: every work item is a walk over an instance's round tables
// (fx_th_assemble_rounds' walk), else one block (fx_th_assemble's); two
// kernels, so that neither path's registers weigh on the other.
template <bool Rounds>
__global__ __launch_bounds__(kThLanes* kThWaves) __attribute__((amdgpu_waves_per_eu(2))) void k_batch_th_assemble(
    const ThAsmArgs* __restrict__ inst, const int2* __restrict__ work, int nwork, int tile)
{
  extern __shared__ double smem[];
  ThTensor& T = *reinterpret_cast<ThTensor*>(smem);
  th_stage_tensor(T, kThLanes * kThWaves);
  __syncthreads();
  // the wave index is uniform across the wave: say so, so that the work item
  // and the instance's arguments are loaded once per wave into scalar registers
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kThLanes);
  const fx::XcdWalk wk = fx::xcd_walk(nwork, kThWaves, w);
  if (Rounds) th_tile_clear(smem + kThTensorLen + w * tile, tile);
  for (int i = wk.first; i < wk.end; i += wk.stride)
  {
    const int2       it = work[i];
    const ThAsmArgs& a  = inst[it.x];
    if (Rounds)
      th_assemble_walk_rt<false>(a, T, smem + kThTensorLen + w * tile, it.y, a.rt_blocks,
                                 a.rt_stride);
    else
      th_assemble_block(it.y, a, T, smem + kThTensorLen + w * tile);
  }
}

__global__ __launch_bounds__(kPatchLanes) __attribute__((amdgpu_waves_per_eu(4))) void k_batch_th_vjp_patch(
    const ThVjpArgs* __restrict__ inst, const PatchArgs* __restrict__ pat,
    const int2* __restrict__ work, int nwork, int max_patch_ldofs)
{
  extern __shared__ __align__(16) unsigned char lds[];
  const fx::XcdWalk wk = fx::xcd_walk(nwork);
  for (int w = wk.first; w < wk.end; w += wk.stride)
  {
    const int2 it = work[w];
    th_vjp_patch_block(it.y, inst[it.x], pat[it.x], kThTables, lds, max_patch_ldofs);
  }
}

__global__ __launch_bounds__(256) void k_batch_vjp_seam(const SeamArgs* __restrict__ inst,
                                                          const int2* __restrict__ work, int nwork)
{
  const fx::XcdWalk wk = fx::xcd_walk(nwork);
  for (int w = wk.first; w < wk.end; w += wk.stride)
  {
    const int2 it = work[w];
    patch_seam_block(it.y, inst[it.x]);
  }
}

// GLS P1/P1 time-step assembly: one wave per 32-node block, as k_gls_assemble.
__global__ __launch_bounds__(64) void k_batch_gls_assemble(const GlsAsmArgs* __restrict__ inst,
                                                          const int2* __restrict__ work, int nwork)
{
  extern __shared__ double tile[];
  const fx::XcdWalk wk = fx::xcd_walk(nwork);
  for (int w = wk.first; w < wk.end; w += wk.stride)
  {
    const int2 it = work[w];
    gls_assemble_block(it.y, inst[it.x], tile);
    wave_sync(); // every lane has stored its tile values: the next block clears the tile
  }
}

// GLS history VJP, both lags, over element patches, and its seam pass.
__global__ __launch_bounds__(kPatchLanes) void k_batch_gls_vjp_patch(
    const GlsVjpArgs* __restrict__ inst, const PatchArgs* __restrict__ pat,
    const GlsPatchOut* __restrict__ out, const int2* __restrict__ work, int nwork,
    int max_patch_ldofs)
{
  extern __shared__ __align__(16) unsigned char lds[];
  const fx::XcdWalk wk = fx::xcd_walk(nwork);
  for (int w = wk.first; w < wk.end; w += wk.stride)
  {
    const int2 it = work[w];
    gls_vjp_patch_block(it.y, inst[it.x], pat[it.x], out[it.x], lds, max_patch_ldofs);
  }
}

__global__ __launch_bounds__(256) void k_batch_gls_vjp_seam(const GlsVjpArgs* __restrict__ inst,
                                                           const SeamArgs* __restrict__ seam,
                                                           const GlsPatchOut* __restrict__ out,
                                                           const int2* __restrict__ work, int nwork)
{
  const fx::XcdWalk wk = fx::xcd_walk(nwork);
  for (int w = wk.first; w < wk.end; w += wk.stride)
  {
    const int2 it = work[w];
    gls_vjp_seam_block(it.y, inst[it.x], seam[it.x], out[it.x]);
  }
}

template <class T>
T* upload(const std::vector<T>& h)
{
  if (h.empty()) return nullptr;
  T* d = nullptr;
  fx::check_hip(hipMalloc(&d, sizeof(T) * h.size()), "fx_batch_create: hipMalloc");
  fx::check_hip(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice),
                "fx_batch_create: hipMemcpy");
  return d;
}

// The six chains of a batch, each on its own stream: P1 assembly, TH
// assembly, P1 VJP, TH VJP, GLS assembly, GLS VJP (each VJP a patch then a
// seam pass).  None reads another's output, so they may run concurrently.
void launch_chain(fx_batch* b, int chain, hipStream_t s)
{
  auto g = [&](int p) { return dim3(b->pgrid[p]); };
  switch (chain)
  {
  case 0:
    if (b->p1_long)
      hipLaunchKernelGGL(k_batch_p1_assemble<true>, g(kP1Asm), dim3(kP1Block), b->p1_lds, s,
                         b->d_p1_asm, b->d_work[kP1Asm], b->nwork[kP1Asm]);
    else
      hipLaunchKernelGGL(k_batch_p1_assemble<false>, g(kP1Asm), dim3(kP1Block), b->p1_lds, s,
                         b->d_p1_asm, b->d_work[kP1Asm], b->nwork[kP1Asm]);
    break;
  case 1:
    if (b->th_rounds)
      hipLaunchKernelGGL(k_batch_th_assemble<true>, g(kThAsm), dim3(kThLanes * kThWaves), b->th_lds,
                         s, b->d_th_asm, b->d_work[kThAsm], b->nwork[kThAsm], b->th_tile);
    else
      hipLaunchKernelGGL(k_batch_th_assemble<false>, g(kThAsm), dim3(kThLanes * kThWaves),
                         b->th_lds, s, b->d_th_asm, b->d_work[kThAsm], b->nwork[kThAsm],
                         b->th_tile);
    break;
  case 2:
    hipLaunchKernelGGL(k_batch_p1_vjp_patch, g(kP1Vjp), dim3(kPatchLanes), b->p1_vjp_lds, s,
                       b->d_p1_vjp, b->d_p1_pat, b->d_work[kP1Vjp], b->nwork[kP1Vjp],
                       b->p1_max_ldofs);
    fx::check_launch("fx_batch: VJP");
    if (b->nwork[kP1VjpSeam] > 0)
      hipLaunchKernelGGL(k_batch_vjp_seam, g(kP1VjpSeam), dim3(256), 0, s, b->d_p1_seam,
                         b->d_work[kP1VjpSeam], b->nwork[kP1VjpSeam]);
    break;
  case 3:
    hipLaunchKernelGGL(k_batch_th_vjp_patch, g(kThVjpPatch), dim3(kPatchLanes), b->th_vjp_lds, s, b->d_th_vjp,
                       b->d_th_pat, b->d_work[kThVjpPatch], b->nwork[kThVjpPatch], b->th_max_ldofs);
    fx::check_launch("fx_batch: VJP");
    if (b->nwork[kThVjpSeam] > 0)
      hipLaunchKernelGGL(k_batch_vjp_seam, g(kThVjpSeam), dim3(256), 0, s, b->d_th_seam,
                         b->d_work[kThVjpSeam], b->nwork[kThVjpSeam]);
    break;
  case 4:
    hipLaunchKernelGGL(k_batch_gls_assemble, g(kGlsAsm), dim3(64), b->gls_lds, s, b->d_gls_asm,
                       b->d_work[kGlsAsm], b->nwork[kGlsAsm]);
    break;
  default:
    hipLaunchKernelGGL(k_batch_gls_vjp_patch, g(kGlsVjpPatch), dim3(kPatchLanes), b->gls_vjp_lds, s,
                       b->d_gls_vjp, b->d_gls_pat, b->d_gls_out, b->d_work[kGlsVjpPatch],
                       b->nwork[kGlsVjpPatch], b->gls_max_ldofs);
    fx::check_launch("fx_batch: VJP");
    if (b->nwork[kGlsVjpSeam] > 0)
      hipLaunchKernelGGL(k_batch_gls_vjp_seam, g(kGlsVjpSeam), dim3(256), 0, s, b->d_gls_vjp,
                         b->d_gls_seam, b->d_gls_out, b->d_work[kGlsVjpSeam], b->nwork[kGlsVjpSeam]);
  }
  fx::check_launch("fx_batch");
}

bool chain_active(const fx_batch* b, int chain, int what)
{
  switch (chain)
  {
  case 0: return (what & FX_BATCH_ASSEMBLE) && b->nwork[kP1Asm] > 0;
  case 1: return (what & FX_BATCH_ASSEMBLE) && b->nwork[kThAsm] > 0;
  case 2: return (what & FX_BATCH_VJP) && b->nwork[kP1Vjp] > 0;
  case 3: return (what & FX_BATCH_VJP) && b->nwork[kThVjpPatch] > 0;
  case 4: return (what & FX_BATCH_ASSEMBLE) && b->nwork[kGlsAsm] > 0;
  default: return (what & FX_BATCH_VJP) && b->nwork[kGlsVjpPatch] > 0;
  }
}

void ensure_branches(fx_batch* b)
{
  if (b->fork) return;
  for (auto& st : b->branch)
    fx::check_hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
  fx::check_hip(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate");
  for (auto& e : b->done)
    fx::check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
}

// Fork the active chains of the batches from stream s onto their branch
// streams and join them back into s (also the shape a capture of s records:
// one graph branch per chain); with FX_BATCH_SERIAL, launch them in order on s.
void launch_batches(fx_batch* const* bs, int count, int what, hipStream_t s)
{
  if (what & FX_BATCH_SERIAL)
  {
    // every chain on s, one after another: no phase shares the chip
    for (int c = 0; c < kChains; ++c)
      for (int i = 0; i < count; ++i)
        if (chain_active(bs[i], c, what)) launch_chain(bs[i], c, s);
    return;
  }
  for (int i = 0; i < count; ++i)
  {
    fx_batch* b = bs[i];
    ensure_branches(b);
    fx::check_hip(hipEventRecord(b->fork, s), "hipEventRecord");
    for (int c = 0; c < kChains; ++c)
    {
      if (!chain_active(b, c, what)) continue;
      fx::check_hip(hipStreamWaitEvent(b->branch[c], b->fork, 0), "hipStreamWaitEvent");
      launch_chain(b, c, b->branch[c]);
      fx::check_hip(hipEventRecord(b->done[c], b->branch[c]), "hipEventRecord");
    }
  }
  for (int i = 0; i < count; ++i)
    for (int c = 0; c < kChains; ++c)
      if (chain_active(bs[i], c, what))
        fx::check_hip(hipStreamWaitEvent(s, bs[i]->done[c], 0), "hipStreamWaitEvent");
}

// Capture launch_batches on `origin` into an instantiated graph; on failure
// the stream leaves capture mode before the error propagates.
hipGraphExec_t capture_batches(fx_batch* const* bs, int count, int what, hipStream_t origin,
                               hipGraph_t* graph_out)
{
  for (int i = 0; i < count; ++i) ensure_branches(bs[i]); // no stream or event creation in capture
  fx::check_hip(hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal),
                "hipStreamBeginCapture");
  try
  {
    launch_batches(bs, count, what, origin);
  }
  catch (...)
  {
    hipGraph_t partial = nullptr;
    if (hipStreamEndCapture(origin, &partial) == hipSuccess && partial)
      (void)hipGraphDestroy(partial);
    throw;
  }
  hipGraph_t graph = nullptr;
  fx::check_hip(hipStreamEndCapture(origin, &graph), "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  const hipError_t st = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (st != hipSuccess)
  {
    (void)hipGraphDestroy(graph);
    fx::check_hip(st, "hipGraphInstantiate");
  }
  *graph_out = graph;
  return exec;
}

} // namespace

extern "C" {

int fx_batch_create(fx_index num_p1, const fx_p1_instance* p1, fx_index num_th,
                    const fx_th_instance* th, fx_index num_gls, const fx_gls_instance* gls,
                    fx_index grid_per_cu, fx_batch** out)
{
  return fx::guard([&] {
    fx::require(out != nullptr && num_p1 >= 0 && num_th >= 0 && num_gls >= 0 && grid_per_cu >= 0
                    && (num_p1 == 0 || p1) && (num_th == 0 || th) && (num_gls == 0 || gls),
                FX_ERR_ARG, "fx_batch_create: invalid arguments");
    *out   = nullptr;
    auto b = std::make_unique<fx_batch>();
    fx::check_hip(hipGetDevice(&b->device), "hipGetDevice");
    int cus = 0;
    fx::check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device),
                  "hipDeviceGetAttribute");
    b->cus    = cus;
    b->grid   = (std::max(1, cus * grid_per_cu) + 7) / 8 * 8; // xcd_walk needs a multiple of 8
    b->num_p1 = num_p1;
    b->num_th = num_th;
    b->num_gls = num_gls;

    std::vector<P1AsmArgs>  pa;
    std::vector<P1VjpArgs>  pv;
    std::vector<ThAsmArgs>  ta;
    std::vector<ThVjpArgs>  tv;
    std::vector<PatchArgs> tp;
    std::vector<SeamArgs>  tsm;
    std::vector<PatchArgs> pp;
    std::vector<SeamArgs>  psm;
    std::vector<GlsAsmArgs>  ga;
    std::vector<GlsVjpArgs>  gv;
    std::vector<PatchArgs>   gp;
    std::vector<SeamArgs>    gsm;
    std::vector<GlsPatchOut> go;
    std::vector<int2>       w[kPhases];
    int                     max_row_len = 1;
    fx_index                max_blk_nnz = 0;
    for (int i = 0; i < num_p1; ++i)
    {
      const fx_p1_instance& q = p1[i];
      fx::require(q.num_nodes > 0 && q.num_elems > 0 && q.max_row_len > 0 && q.max_row_len <= 255,
                  FX_ERR_ARG, "fx_batch_create: invalid P1 instance sizes");
      fx::require(fx::fits32(q.num_nodes, 8) && fx::fits32(q.num_elems, 8), FX_ERR_OVERFLOW,
                  "fx_batch_create: P1 instance arrays beyond 4 GiB");
      fx::require(q.x && q.y && q.conn && q.inc_off && q.inc && q.row_ptr && q.res && q.vals,
                  FX_ERR_ARG, "fx_batch_create: P1 instance lacks an array");
      // rows past the tile width add into vals in global memory (fx_p1_dev.hpp)
      max_row_len = std::max(max_row_len, std::min(static_cast<int>(q.max_row_len), kP1MaxRowLen));
      if (q.max_row_len > kP1MaxRowLen) b->p1_long = true;
      pa.push_back({q.num_nodes, q.x, q.y, q.conn, q.m, 1, q.f, q.u, q.inc_off,
                    reinterpret_cast<const int2*>(q.inc), q.row_ptr, q.vals, q.res, 0});
      pv.push_back({q.num_nodes, q.x, q.y, q.conn, q.m, q.u, q.lam, q.bc_mask, q.inc_off,
                    reinterpret_cast<const int2*>(q.inc), q.jt, q.g});
      const int nb = static_cast<int>(fx::blocks_for(q.num_nodes, kP1Block));
      for (int k = 0; k < nb; ++k) w[kP1Asm].push_back(make_int2(i, k));
      const fx_patch_plan& pl = q.vjp;
      fx::require(pl.num_patches >= 0 && pl.num_seam >= 0 && pl.num_parts >= 0
                      && pl.max_patch_ldofs >= 0 && pl.max_patch_ldofs <= 3 * kPatchLanes
                      && fx::fits32(pl.num_parts, 8),
                  FX_ERR_ARG, "fx_batch_create: invalid P1 VJP patch plan");
      const bool vjp = q.lam && q.jt && q.g && q.m && q.u && pl.num_patches > 0 && pl.patch_off
                       && pl.patch_elem && pl.elem_ldof && pl.ldof_off && pl.ldof_dof
                       && pl.ldof_part && pl.ldof_end && pl.seam_off
                       && (pl.num_seam == 0 || pl.seam_dof) && (pl.num_parts == 0 || pl.parts);
      pp.push_back({pl.patch_off, pl.patch_elem, pl.elem_vert, pl.elem_ldof, pl.ldof_off,
                    pl.ldof_dof, pl.ldof_part, pl.ldof_end, q.jt, pl.parts});
      psm.push_back({pl.num_seam, pl.seam_off, pl.seam_dof, pl.parts, q.lam, q.bc_mask, q.jt});
      if (vjp)
      {
        b->p1_max_ldofs = std::max(b->p1_max_ldofs, static_cast<int>(pl.max_patch_ldofs));
        const int ns = static_cast<int>(fx::blocks_for(pl.num_seam, 256));
        for (int k = 0; k < pl.num_patches; ++k) w[kP1Vjp].push_back(make_int2(i, k));
        for (int k = 0; k < ns; ++k) w[kP1VjpSeam].push_back(make_int2(i, k));
      }
    }
    const int acc_cap = kP1Block * max_row_len;
    for (auto& a : pa) a.acc_cap = acc_cap;
    b->p1_lds = static_cast<std::size_t>(acc_cap) * 8 + kIncStage * 8;
    b->th_rounds = num_th > 0;
    for (int i = 0; i < num_th; ++i)
      b->th_rounds = b->th_rounds && th[i].rt_off && th[i].rt_elem && th[i].rt_meta && th[i].blk_lanes;
    for (int i = 0; i < num_th; ++i)
    {
      const fx_th_instance& q = th[i];
      fx::require(q.ndof > 0 && q.num_elems > 0 && q.num_blocks > 0 && q.max_block_nnz >= 0,
                  FX_ERR_ARG, "fx_batch_create: invalid Taylor-Hood instance sizes");
      fx::require(fx::fits32(q.ndof, 8) && fx::fits32(q.num_elems, 8)
                      && fx::fits32(15 * static_cast<std::int64_t>(q.num_elems), 16),
                  FX_ERR_OVERFLOW, "fx_batch_create: Taylor-Hood instance arrays beyond 4 GiB");
      fx::require(q.xv && q.yv && q.vconn3 && q.dofs15 && q.state && q.inc_off && q.inc_elem
                      && q.inc_meta && q.blk_row && q.blk_flags && q.row_ptr && q.vals && q.res,
                  FX_ERR_ARG, "fx_batch_create: Taylor-Hood instance lacks an array");
      max_blk_nnz = std::max(max_blk_nnz, q.max_block_nnz);
      // with round tables (on every instance) a work item is a walk over
      // kThRtBatchWalk blocks of the instance (k, k + M, ... < num_blocks, M
      // the number of walks), so that walks next to each other take
      // neighbouring blocks
      const bool rounds = b->th_rounds;
      const int  walks  = rounds ? (q.num_blocks + kThRtBatchWalk - 1) / kThRtBatchWalk : 0;
      ta.push_back({q.xv, q.yv, q.vconn3, q.dofs15, q.state, q.nu, q.nu_e, q.inc_off, q.inc_elem,
                    reinterpret_cast<const uint4*>(q.inc_meta), q.blk_row, q.blk_flags, q.blk_lanes,
                    q.row_ptr, q.vals, q.res, rounds ? q.rt_off : nullptr,
                    rounds ? q.rt_elem : nullptr,
                    rounds ? reinterpret_cast<const uint4*>(q.rt_meta) : nullptr,
                    static_cast<int>(q.num_blocks), walks});
      for (int k = 0; k < (rounds ? walks : q.num_blocks); ++k) w[kThAsm].push_back(make_int2(i, k));
      const fx_patch_plan& pv = q.vjp;
      const bool vjp = q.lam && q.jt && pv.num_patches > 0 && pv.patch_off && pv.patch_elem
                       && pv.elem_vert && pv.elem_ldof && pv.ldof_off && pv.ldof_dof && pv.ldof_part && pv.ldof_end
                       && pv.seam_off && (pv.num_seam == 0 || pv.seam_dof)
                       && (pv.num_parts == 0 || pv.parts);
      fx::require(pv.num_patches >= 0 && pv.num_seam >= 0 && pv.num_parts >= 0
                      && pv.max_patch_ldofs >= 0 && pv.max_patch_ldofs <= 15 * kPatchLanes
                      && fx::fits32(pv.num_parts, 8),
                  FX_ERR_ARG, "fx_batch_create: invalid Taylor-Hood VJP patch plan");
      tv.push_back({q.num_elems, q.xv, q.yv, q.vconn3, q.dofs15, q.state, q.nu, q.nu_e, q.lam,
                    q.bc_mask, nullptr, nullptr, q.g});
      tp.push_back({pv.patch_off, pv.patch_elem, pv.elem_vert, pv.elem_ldof, pv.ldof_off,
                    pv.ldof_dof, pv.ldof_part, pv.ldof_end, q.jt, pv.parts});
      tsm.push_back({pv.num_seam, pv.seam_off, pv.seam_dof, pv.parts, q.lam, q.bc_mask, q.jt});
      if (vjp) b->th_max_ldofs = std::max(b->th_max_ldofs, static_cast<int>(pv.max_patch_ldofs));
      if (vjp)
      {
        const int ns = static_cast<int>(fx::blocks_for(pv.num_seam, 256));
        for (int k = 0; k < pv.num_patches; ++k) w[kThVjpPatch].push_back(make_int2(i, k));
        for (int k = 0; k < ns; ++k) w[kThVjpSeam].push_back(make_int2(i, k));
      }
    }
    int gls_tile = 0;
    for (int i = 0; i < num_gls; ++i)
    {
      const fx_gls_instance& q = gls[i];
      // fx_gls_assemble's and fx_gls_hist_vjp_patch's checks
      // (NavierStokesModel.cpp:35-46, Assembly.hpp:86-104)
      fx::require(q.num_nodes > 0 && q.num_states > 0 && q.num_elems > 0 && q.max_tile >= 0
                      && q.num_hist > 0,
                  FX_ERR_ARG, "fx_batch_create: invalid GLS instance sizes");
      fx::require(q.dt > 0.0 && std::isfinite(q.dt) && q.rho > 0.0 && std::isfinite(q.rho)
                      && q.mu > 0.0 && std::isfinite(q.mu),
                  FX_ERR_ARG, "fx_batch_create: NavierStokesModel requires positive finite "
                              "time step, density and viscosity");
      fx::require(fx::fits32(q.num_states, 8 * q.num_hist) && fx::fits32(q.num_nodes, 12),
                  FX_ERR_OVERFLOW, "fx_batch_create: GLS instance arrays beyond 4 GiB");
      fx::require(q.x && q.y && q.conn3 && q.dofs9 && q.node_rows && q.hist && q.nxt && q.inc_off
                      && q.inc_elem && q.inc_meta && q.blk_runs && q.node_tile && q.res,
                  FX_ERR_ARG, "fx_batch_create: GLS instance lacks an array");
      fx::require(!q.vals || q.max_tile <= kGlsMaxTile, FX_ERR_UNSUPPORTED,
                  "fx_batch_create: a GLS block's rows exceed the 64 KiB LDS tile");
      if (q.vals) gls_tile = std::max(gls_tile, static_cast<int>(q.max_tile > 0 ? q.max_tile : 1));
      const GlsConst c = gls_const(q.rho, q.mu, q.dt);
      ga.push_back({static_cast<int>(q.num_nodes), static_cast<int>(q.num_states), q.x, q.y,
                    q.conn3, q.dofs9, q.node_rows, static_cast<int>(q.step),
                    static_cast<int>(q.num_hist), q.hist, q.nxt, c, q.inc_off, q.inc_elem,
                    reinterpret_cast<const uint4*>(q.inc_meta),
                    reinterpret_cast<const int4*>(q.blk_runs), q.node_tile, q.vals, q.res,
                    static_cast<int>(q.max_tile)});
      const int nb = static_cast<int>(fx::blocks_for(q.num_nodes, kGlsBlock));
      for (int k = 0; k < nb; ++k) w[kGlsAsm].push_back(make_int2(i, k));
      const fx_patch_plan& pl = q.vjp;
      fx::require(pl.num_patches >= 0 && pl.num_seam >= 0 && pl.num_parts >= 0
                      && pl.max_patch_ldofs >= 0 && pl.max_patch_ldofs <= 3 * kPatchLanes
                      && fx::fits32(pl.num_parts, 32),
                  FX_ERR_ARG, "fx_batch_create: invalid GLS VJP patch plan");
      fx::require(!q.out_lag1 || q.num_hist > 1, FX_ERR_ARG,
                  "fx_batch_create: Time assembly variable block is invalid");
      const bool vjp = q.adj && (q.out_lag0 || q.out_lag1) && pl.num_patches > 0 && pl.patch_off
                       && pl.patch_elem && pl.elem_ldof && pl.ldof_off && pl.ldof_dof
                       && pl.ldof_part && pl.ldof_end && pl.seam_off
                       && (pl.num_seam == 0 || pl.seam_dof) && (pl.num_parts == 0 || pl.parts);
      gv.push_back({static_cast<int>(q.num_nodes), static_cast<int>(q.num_states), q.x, q.y,
                    nullptr, nullptr, q.node_rows, static_cast<int>(q.step),
                    static_cast<int>(q.num_hist), 0, q.hist, q.nxt, q.adj, c, nullptr, nullptr,
                    nullptr, nullptr, static_cast<int>(q.num_elems)});
      gp.push_back({pl.patch_off, pl.patch_elem, pl.elem_vert, pl.elem_ldof, pl.ldof_off,
                    pl.ldof_dof, pl.ldof_part, pl.ldof_end, nullptr, pl.parts});
      gsm.push_back({pl.num_seam, pl.seam_off, pl.seam_dof, pl.parts, nullptr, nullptr, nullptr});
      go.push_back({q.out_lag0, q.out_lag1});
      if (vjp)
      {
        b->gls_max_ldofs = std::max(b->gls_max_ldofs, static_cast<int>(pl.max_patch_ldofs));
        const int ns = static_cast<int>(fx::blocks_for(pl.num_seam, 256));
        for (int k = 0; k < pl.num_patches; ++k) w[kGlsVjpPatch].push_back(make_int2(i, k));
        for (int k = 0; k < ns; ++k) w[kGlsVjpSeam].push_back(make_int2(i, k));
      }
    }
    b->gls_lds     = static_cast<std::size_t>(gls_tile) * 8;
    b->gls_vjp_lds = gls_patch_lds(b->gls_max_ldofs);
    b->th_tile = static_cast<int>(max_blk_nnz);
    b->th_lds  = sizeof(ThTensor) + static_cast<std::size_t>(max_blk_nnz) * 8 * kThWaves;
    fx::require(b->p1_lds <= 160 * 1024 && b->th_lds <= 160 * 1024, FX_ERR_UNSUPPORTED,
                "fx_batch_create: LDS tile exceeds the workgroup limit");
    // Persistent grids (grid_per_cu > 0) cover at least every workgroup slot
    // the phase kernel can hold resident, so no phase runs below its
    // occupancy; grid_per_cu = 0 gives one workgroup per work item (per four
    // for the Taylor-Hood assembly, whose waves walk the list on their own),
    // so that concurrent phases share the chip as their workgroups retire.
    auto occ = [&](int p, const void* k, int threads, std::size_t lds) {
      if (grid_per_cu == 0)
      {
        const std::int64_t items = w[p].size(), per = p == kThAsm ? kThWaves : 1;
        return static_cast<int>(std::max<std::int64_t>(8, (items + per - 1) / per + 7) / 8 * 8);
      }
      int n = 0;
      fx::check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, threads, lds),
                    "hipOccupancyMaxActiveBlocksPerMultiprocessor");
      return (std::max(b->grid, cus * n) + 7) / 8 * 8;
    };
    b->pgrid[kP1Asm] =
        occ(kP1Asm,
            b->p1_long ? reinterpret_cast<const void*>(k_batch_p1_assemble<true>)
                       : reinterpret_cast<const void*>(k_batch_p1_assemble<false>),
            kP1Block, b->p1_lds);
    b->pgrid[kThAsm] = occ(kThAsm,
                           b->th_rounds ? reinterpret_cast<const void*>(k_batch_th_assemble<true>)
                                        : reinterpret_cast<const void*>(k_batch_th_assemble<false>),
                           kThLanes * kThWaves, b->th_lds);
    b->p1_vjp_lds    = p1_patch_lds(b->p1_max_ldofs);
    b->pgrid[kP1Vjp] = occ(kP1Vjp, reinterpret_cast<const void*>(k_batch_p1_vjp_patch), kPatchLanes,
                           b->p1_vjp_lds);
    b->pgrid[kP1VjpSeam] =
        occ(kP1VjpSeam, reinterpret_cast<const void*>(k_batch_vjp_seam), 256, 0);
    b->th_vjp_lds = th_patch_lds(b->th_max_ldofs);
    b->pgrid[kThVjpPatch] =
        occ(kThVjpPatch, reinterpret_cast<const void*>(k_batch_th_vjp_patch), kPatchLanes,
            b->th_vjp_lds);
    b->pgrid[kThVjpSeam] =
        occ(kThVjpSeam, reinterpret_cast<const void*>(k_batch_vjp_seam), 256, 0);
    b->pgrid[kGlsAsm] =
        occ(kGlsAsm, reinterpret_cast<const void*>(k_batch_gls_assemble), 64, b->gls_lds);
    b->pgrid[kGlsVjpPatch] = occ(kGlsVjpPatch, reinterpret_cast<const void*>(k_batch_gls_vjp_patch),
                                 kPatchLanes, b->gls_vjp_lds);
    b->pgrid[kGlsVjpSeam] =
        occ(kGlsVjpSeam, reinterpret_cast<const void*>(k_batch_gls_vjp_seam), 256, 0);
    b->d_p1_asm = upload(pa);
    b->d_p1_vjp = upload(pv);
    b->d_th_asm = upload(ta);
    b->d_th_vjp = upload(tv);
    b->d_th_pat  = upload(tp);
    b->d_th_seam = upload(tsm);
    b->d_p1_pat  = upload(pp);
    b->d_p1_seam = upload(psm);
    b->d_gls_asm  = upload(ga);
    b->d_gls_vjp  = upload(gv);
    b->d_gls_pat  = upload(gp);
    b->d_gls_seam = upload(gsm);
    b->d_gls_out  = upload(go);
    for (int p = 0; p < kPhases; ++p)
    {
      b->nwork[p]  = static_cast<int>(w[p].size());
      b->d_work[p] = upload(w[p]);
    }
    *out = b.release();
  });
}

int fx_batch_work(const fx_batch* b, fx_index* counts9, fx_index* grid)
{
  return fx::guard([&] {
    fx::require(b && counts9, FX_ERR_ARG, "fx_batch_work: null argument");
    for (int p = 0; p < kPhases; ++p) counts9[p] = b->nwork[p];
    if (grid) *grid = b->grid;
  });
}

int fx_batch_run(fx_batch* b, int what, void* stream)
{
  return fx::guard([&] {
    fx::require(b != nullptr, FX_ERR_ARG, "fx_batch_run: null batch");
    launch_batches(&b, 1, what, fx::as_stream(stream));
  });
}

int fx_batch_capture(fx_batch* b, int what)
{
  return fx::guard([&] {
    fx::require(b != nullptr, FX_ERR_ARG, "fx_batch_capture: null batch");
    if (b->exec)
    {
      (void)hipGraphExecDestroy(b->exec);
      b->exec = nullptr;
    }
    if (b->graph)
    {
      (void)hipGraphDestroy(b->graph);
      b->graph = nullptr;
    }
    if (!b->cap_stream)
      fx::check_hip(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking), "hipStreamCreate");
    b->exec       = capture_batches(&b, 1, what, b->cap_stream, &b->graph);
    b->graph_what = what;
  });
}

int fx_batch_launch(fx_batch* b, void* stream)
{
  return fx::guard([&] {
    fx::require(b && b->exec, FX_ERR_ARG, "fx_batch_launch: capture the batch first");
    fx::check_hip(hipGraphLaunch(b->exec, fx::as_stream(stream)), "hipGraphLaunch");
  });
}

int fx_batch_graph_create(fx_index count, fx_batch* const* batches, int what,
                          fx_batch_graph** out)
{
  return fx::guard([&] {
    fx::require(out && count > 0 && batches, FX_ERR_ARG, "fx_batch_graph_create: invalid arguments");
    *out = nullptr;
    for (fx_index i = 0; i < count; ++i)
    {
      fx::require(batches[i] != nullptr, FX_ERR_ARG, "fx_batch_graph_create: null batch");
      for (fx_index j = 0; j < i; ++j)
        fx::require(batches[j] != batches[i], FX_ERR_ARG,
                    "fx_batch_graph_create: a batch is listed twice");
    }
    auto g = std::make_unique<fx_batch_graph>();
    fx::check_hip(hipStreamCreateWithFlags(&g->origin, hipStreamNonBlocking), "hipStreamCreate");
    g->exec = capture_batches(batches, static_cast<int>(count), what, g->origin, &g->graph);
    *out    = g.release();
  });
}

int fx_batch_graph_launch(fx_batch_graph* g, void* stream)
{
  return fx::guard([&] {
    fx::require(g && g->exec, FX_ERR_ARG, "fx_batch_graph_launch: null graph");
    fx::check_hip(hipGraphLaunch(g->exec, fx::as_stream(stream)), "hipGraphLaunch");
  });
}

void fx_batch_graph_destroy(fx_batch_graph* g) { delete g; }

void fx_batch_destroy(fx_batch* b) { delete b; }

} // extern "C"
