// fhd_launch.h -- one launch path for every step kernel, with a batched form.
//
// Each step kernel is a device body that takes its argument struct and a block descriptor
// (Blk), wrapped twice:
//   * the single form, k_x(A a), runs one context's launch (block = blockIdx);
//   * the batched form, kb_x(table), runs the same stage of many contexts in one dispatch.
//     Workgroup w reads table.blk[w] = {entry, bx, by, bz}; entry e has its own grid
//     (table.dims[e]) and its own argument struct at table.args + e * table.stride.
// A workgroup of the batched form executes exactly the instructions of the matching
// workgroup of the single form, so a batched context gets the same bits as a solo one.
//
// Host side, fhd::launch() either launches the single form or, while a thread-local
// Recorder is installed, appends the launch (family, batched entry point, grid, argument
// bytes) to the recorder.  The batch code merges recorded launches of many contexts into
// batched dispatches.
#pragma once
#include <hip/hip_runtime.h>
#include <string.h>
#include <stdexcept>
#include <vector>

namespace fhd {

struct Blk {
  int x, y, z;     // block index
  int nx, ny, nz;  // grid extent
};

#define FHD_BODY __device__ __forceinline__

FHD_BODY Blk blk_here() {
  return Blk{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x, (int)gridDim.y, (int)gridDim.z};
}

// Parameters of every batched kernel.  The pointers are restrict so that the argument loads
// (uniform per workgroup) stay scalar loads.
#define FHD_BATCH_PARAMS                                                                           \
  const int4 *__restrict__ fhd_blk, const int4 *__restrict__ fhd_dims, const unsigned char *__restrict__ fhd_args, \
      int fhd_stride

template <class A>
FHD_BODY const A& batch_entry(const int4* __restrict__ blk, const int4* __restrict__ dims,
                              const unsigned char* __restrict__ args, int stride, Blk& b) {
  const int4 e = blk[blockIdx.x];
  const int4 n = dims[e.x];
  b = Blk{e.y, e.z, e.w, n.x, n.y, n.z};
  return *reinterpret_cast<const A*>(args + (size_t)e.x * (size_t)stride);
}

// Defines the single form `name` and the batched form `name##_b` of a kernel.  TPL is the
// parenthesised template header (`()` for none); the trailing arguments are the body call,
// which sees the argument struct as `a` and the block descriptor as `b`.
#define FHD_UNPAREN(x) FHD_UNPAREN_ x
#define FHD_UNPAREN_(...) __VA_ARGS__
#define FHD_KERNEL_PAIR(TPL, NT, name, A, ...)                                                    \
  FHD_UNPAREN(TPL) __global__ __launch_bounds__(NT) void name(A a) {                               \
    const Blk b = blk_here();                                                                      \
    __VA_ARGS__;                                                                                   \
  }                                                                                                \
  FHD_UNPAREN(TPL) __global__ __launch_bounds__(NT) void name##_b(FHD_BATCH_PARAMS) {             \
    Blk b;                                                                                         \
    const A& a = batch_entry<A>(fhd_blk, fhd_dims, fhd_args, fhd_stride, b);                       \
    __VA_ARGS__;                                                                                   \
  }

// The same pair with a minimum of WPE waves per SIMD (a register budget of 512 / WPE VGPRs).
#define FHD_KERNEL_PAIR_OCC(TPL, NT, WPE, name, A, ...)                                           \
  FHD_UNPAREN(TPL) __global__ __launch_bounds__(NT, WPE) void name(A a) {                          \
    const Blk b = blk_here();                                                                      \
    __VA_ARGS__;                                                                                   \
  }                                                                                                \
  FHD_UNPAREN(TPL) __global__ __launch_bounds__(NT, WPE) void name##_b(FHD_BATCH_PARAMS) {        \
    Blk b;                                                                                         \
    const A& a = batch_entry<A>(fhd_blk, fhd_dims, fhd_args, fhd_stride, b);                       \
    __VA_ARGS__;                                                                                   \
  }

// Stage families: launches of one family from different contexts merge into one dispatch
// (one per distinct kernel instance).
enum Family {
  FAM_BIN = 1, FAM_SCAN, FAM_SCATTER, FAM_TILE_SORT, FAM_WCA, FAM_ZERO, FAM_SRC_PREP, FAM_SPREAD,
  FAM_FLUID_W, FAM_FFT_XY_FWD, FAM_FFT_Z, FAM_FFT_XY_INV, FAM_INTERP, FAM_STEP, FAM_BONDED,
  FAM_PVEC, FAM_LAG2, FAM_MASS_FLUID, FAM_WCA_LIST, FAM_SORT_ONE
};

using BatchedFn = void (*)(const int4*, const int4*, const unsigned char*, int);

struct LaunchRecord {
  int family;
  const void* fn;  // batched entry point
  dim3 grid;
  int threads;
  std::vector<unsigned char> args;
};

struct Recorder {
  std::vector<LaunchRecord> list;
};

// Installed by the batch code while it records a context's step on this thread.
extern thread_local Recorder* t_recorder;

template <class A>
void launch(void (*single)(A), BatchedFn batched, int family, dim3 grid, int threads, const A& a, hipStream_t s) {
  if ((size_t)grid.x * grid.y * grid.z == 0) return;
  if (Recorder* r = t_recorder) {
    LaunchRecord rec;
    rec.family = family;
    rec.fn = reinterpret_cast<const void*>(batched);
    rec.grid = grid;
    rec.threads = threads;
    rec.args.resize(sizeof(A));
    memcpy(rec.args.data(), &a, sizeof(A));
    r->list.push_back(std::move(rec));
    return;
  }
  hipLaunchKernelGGL(single, grid, dim3(threads), 0, s, a);
}

}  // namespace fhd
