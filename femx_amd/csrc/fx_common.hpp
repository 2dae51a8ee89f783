// Shared helpers of the femx_amd C-ABI: status codes, error text, argument
// and alias checks, and HIP launch plumbing.  No exception crosses the C-ABI:
// every entry point runs its body through fx::guard(), which turns an
// fx::Error into a status code and keeps the message for fx_last_error().
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/femx_amd.h"

namespace fx
{

struct Error : std::runtime_error
{
  int code;
  Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

void set_error(const std::string& msg);

template <class F>
int guard(F&& body)
{
  try
  {
    body();
    return FX_OK;
  }
  catch (const Error& e)
  {
    set_error(e.what());
    return e.code;
  }
  catch (const std::bad_alloc&)
  {
    set_error("femx_amd: host allocation failed");
    return FX_ERR_ALLOC;
  }
  catch (const std::exception& e)
  {
    set_error(e.what());
    return FX_ERR_INTERNAL;
  }
}

inline void require(bool cond, int code, const char* msg)
{
  if (!cond) throw Error(code, msg);
}

inline void check_hip(hipError_t st, const char* what)
{
  if (st != hipSuccess)
    throw Error(FX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(st));
}

// After a launch: report launch-configuration errors at the call site.
inline void check_launch(const char* what) { check_hip(hipGetLastError(), what); }

// Byte ranges [a, a+na) and [b, b+nb) overlap.
inline bool overlaps(const void* a, std::size_t na, const void* b, std::size_t nb)
{
  if (a == nullptr || b == nullptr || na == 0 || nb == 0) return false;
  auto pa = reinterpret_cast<std::uintptr_t>(a);
  auto pb = reinterpret_cast<std::uintptr_t>(b);
  return pa < pb + nb && pb < pa + na;
}

// Blocks of consecutive rows land on the same XCD in order: hardware deals
// workgroups to the 8 XCDs round-robin, so remap the dispatch index to keep
// neighbouring row blocks (which share elements and nodes) in one L2.
__device__ __forceinline__ int xcd_block(int nblocks)
{
  const int bid = blockIdx.x;
  const int per = nblocks / 8, rem = nblocks % 8;
  const int xcd = bid % 8, k = bid / 8;
  // XCD x owns logical blocks [x*per + min(x,rem), ...) of length per + (x<rem),
  // and receives exactly that many dispatch indices (those with bid % 8 == x).
  return xcd * per + min(xcd, rem) + k;
}

// Persistent walk over n work items by `sub` walkers per workgroup (walker
// j): the walkers on each XCD take one contiguous eighth of the list, so
// neighbouring items -- which share mesh data -- meet in one L2.  Requires
// gridDim.x to be a multiple of 8 (the batch rounds its grid up to one).
struct XcdWalk
{
  int first, end, stride;
};

__device__ __forceinline__ XcdWalk xcd_walk(int n, int sub = 1, int j = 0)
{
  const int xcd = blockIdx.x % 8, k = blockIdx.x / 8, per = gridDim.x / 8;
  const int chunk = (n + 7) / 8;
  const int begin = xcd * chunk;
  return {begin + k * sub + j, min(n, begin + chunk), per * sub};
}

// n items of `bytes` each fit under 4 GiB: the kernels gather such arrays
// through 32-bit byte offsets.
inline bool fits32(std::int64_t n, int bytes) { return n >= 0 && n * bytes <= 0xffffffffLL; }

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

inline unsigned blocks_for(std::int64_t n, int block)
{
  return static_cast<unsigned>((n + block - 1) / block);
}

} // namespace fx
