// Recording and batched replay of the field-stage launches.
//
// A batch of image pairs runs the field stages of one pass level (predictor
// interpolation, validation, replacement, smoothn and the counter resets)
// for every pair.  Launched pair by pair these are dozens of tiny launches
// per pair.  Instead, the per-pair launch sequence is recorded once -- each
// launch site appends (kind, argument struct, grid) to the recording rather
// than launching -- and launch j of every pair with the same sequence then
// runs as ONE launch of kind j: a 1-D grid over all the pairs' blocks, each
// block finding its pair by binary search over the pairs' first blocks and
// running the same kernel body with the pair's arguments and its virtual
// block index.  The bodies are shared with the single-pair kernels, so the
// results are identical.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <vector>

namespace piv {

// Virtual block index and grid of one body invocation.
struct VB {
  int bx, by, bz;
  int gx, gy, gz;
};

__device__ __forceinline__ VB vb_native() {
  return VB{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x, (int)gridDim.y, (int)gridDim.z};
}

// grids[i] = (gx, gy, gz, first block of item i), first blocks ascending.
__device__ __forceinline__ int batch_item(const int4* grids, int n, int block) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (grids[mid].w <= block) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ VB vb_item(int4 g, int block) {
  const int local = block - g.w;
  const int xy = g.x * g.y;
  return VB{local % g.x, (local / g.x) % g.y, local / xy, g.x, g.y, g.z};
}

// Kinds of recordable launches (every kernel of the FP32 field stages).
enum class LaunchKind : int32_t {
  kFill = 0,
  kCopy,
  kScale,
  kGridInterp,
  kValidate,
  kValidateWide,
  kFieldStats,
  kSpring,
  kNbrReplace,
  kNbrReplaceWide,
  kReset,
  kGemm,
  kDct2Gcv,
  kGcvGrid,
  kNewton1,
  kNewton2,
  kGcvFinal,
};

struct RecordedLaunch {
  LaunchKind kind;
  int block;          // threads per block
  int gx, gy, gz;     // grid of the single launch
  std::vector<unsigned char> args;
};

// While non-null, the recordable launch sites append to this list and
// launch nothing (set only by the thread building a batch).
extern thread_local std::vector<RecordedLaunch>* g_launch_record;

template <typename A>
inline bool record_launch(LaunchKind kind, const A& a, dim3 grid, int block) {
  if (!g_launch_record) return false;
  RecordedLaunch r;
  r.kind = kind;
  r.block = block;
  r.gx = (int)grid.x;
  r.gy = (int)grid.y;
  r.gz = (int)grid.z;
  r.args.resize(sizeof(A));
  std::memcpy(r.args.data(), &a, sizeof(A));
  g_launch_record->push_back(std::move(r));
  return true;
}

// One batched launch of `kind` over n items: args is a device array of the
// items' argument structs, grids their (gx, gy, gz, first block);
// total_blocks the sum of their blocks.  launch_field_batch takes the kinds
// of the field kernels, launch_smoothn_batch those of the DCT / GCV kernels
// (kGemm and after).
hipError_t launch_field_batch(LaunchKind kind, const void* args, const int4* grids, int n, int total_blocks,
                              hipStream_t stream);
hipError_t launch_smoothn_batch(LaunchKind kind, const void* args, const int4* grids, int n, int total_blocks,
                                hipStream_t stream);

// Batched wrapper: Body::run(const Args&, VB) is the kernel body.
template <typename Body, int THREADS>
__global__ __launch_bounds__(THREADS) void batched_kernel(const typename Body::Args* args, const int4* grids, int n) {
  const int i = batch_item(grids, n, (int)blockIdx.x);
  const typename Body::Args a = args[i];
  Body::run(a, vb_item(grids[i], (int)blockIdx.x));
}

template <typename Body, int THREADS>
hipError_t launch_batched(const void* args, const int4* grids, int n, int total_blocks, hipStream_t stream) {
  if (n <= 0 || total_blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL((batched_kernel<Body, THREADS>), dim3(total_blocks), dim3(THREADS), 0, stream,
                     reinterpret_cast<const typename Body::Args*>(args), grids, n);
  return hipGetLastError();
}

}  // namespace piv
