// dlbc_hip.hip -- context, data movement, reductions and the C ABI of
// include/dlbc_hip.h.  The time-step kernels live in kernels_*.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dlbc_hip.h"
#include "kernels_batch.hpp"
#include "kernels_fused.hpp"
#include "kernels_init.hpp"
#include "kernels_march.hpp"
#include "kernels_shifted.hpp"
#include "kernels_split.hpp"
#include "kernels_twopass.hpp"

using namespace dlbc;

namespace {

constexpr int kMaxC = 4;
constexpr int kMaxKernels = 8;

}  // namespace

struct dlbc_gpu_ctx {
  int d = 0, q = 0, C = 0, storage = 0, device = 0;
  int n[3] = {1, 1, 1};
  Geo geo{};
  hipStream_t stream = nullptr;
  void *popA = nullptr, *popB = nullptr;  // populations (double / float / __half)
  int elem = 8;                            // bytes per stored population
  double rho0[kMaxC] = {1.0, 1.0, 1.0, 1.0};
  ShiftParams sparams{};
  double *S = nullptr;                      // psi (fluid) / rho (solid)
  float *P = nullptr;                       // shifted storage: psi in FP32 (split schedule)
  uint8_t *mask = nullptr;
  uint32_t *nbw = nullptr;  // solid-neighbour words (bit i: x + c_i solid)
  int has_solid = 0;
  bool params_set = false, pops_uploaded[kMaxC] = {false, false, false, false};
  bool solid_rho_dirty = true;
  StepParams p{};
  int schedule = DLBC_SCHED_AUTO;
  int target_blocks = 0;  // fused kernels: minimum workgroups per launch (0: default)
  int march_rows = 0;     // march schedule: z-rows per workgroup (0: automatic)
  int march_hold = -1;    // march schedule: populations held in registers (-1: automatic)
  // shifted storage: the tiled kernel over this lattice alone (a batch of one)
  void *bufX = nullptr, *bufY = nullptr;  // the two population buffers as allocated
  BatchLattice *d_self = nullptr;
  BatchLattice self_tab{};
  bool self_valid = false;
  BatchItem *d_items = nullptr;
  int nitems = 0, items_target = -1;
  std::string err;
  int64_t dev_bytes = 0;
  hipEvent_t ev[2 * kMaxKernels + 2] = {};
  bool events_ready = false;
  // scratch for reductions and transfers
  double *scratch = nullptr;
  size_t scratch_bytes = 0;
  // dlbc_gpu_step_graph: two captured steps per buffer parity, and the host
  // state their kernel arguments were captured from
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  std::vector<unsigned char> graph_key;
};

namespace {

int fail(dlbc_gpu_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(ctx, call)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail((ctx), DLBC_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));      \
  } while (0)

int ensure_scratch(dlbc_gpu_ctx *c, size_t bytes) {
  if (c->scratch_bytes >= bytes) return DLBC_OK;
  if (c->scratch) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipFree(c->scratch));
    c->dev_bytes -= (int64_t)c->scratch_bytes;
    c->scratch = nullptr;
    c->scratch_bytes = 0;
  }
  if (hipMalloc(&c->scratch, bytes) != hipSuccess)
    return fail(c, DLBC_ERR_NOMEM, "cannot allocate %zu bytes of device scratch", bytes);
  c->scratch_bytes = bytes;
  c->dev_bytes += (int64_t)bytes;
  return DLBC_OK;
}

inline int blocks_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

// Block order of the split kernels (split_block): chunk -1 picks it from the
// geometry, 0 keeps site order, K > 0 gives the XCDs K blocks per turn.  The
// automatic choice is site order (see kSplitAutoChunk).
void apply_block_order(dlbc_gpu_ctx *c, int chunk) {
  const int spb = c->storage == DLBC_STORAGE_F64 ? kSplitThreads / 2 : kShiftSites;
  if (chunk < 0) {
    chunk = 0;
    const int64_t plane = (int64_t)c->n[1] * c->n[2];
    if (kSplitAutoChunk && plane % spb == 0 && (plane / spb) % 8 == 0)
      chunk = (int)(plane / spb / 8);  // an eighth of an x-plane per turn
  }
  c->geo.order_chunk = chunk;
  c->geo.order_main = split_order_main((int)((c->geo.sites + spb - 1) / spb), chunk);
}

// ---- population loaders -----------------------------------------------------
// A loader returns population i of component f at site s in FP64: the stored
// value (FP64 storage) or w_i rho0_f + the stored deviation (shifted storage).
struct LoadF64 {
  const double *A;
  int64_t sp;
  int q;
  __device__ double operator()(int f, int i, int s) const {
    return A[(int64_t)(f * q + i) * sp + s];
  }
};
template <class St>
struct LoadShift {
  const St *A;
  int64_t sp;
  double rho0[4];
  __device__ double operator()(int f, int i, int s) const {
    return D3Q19::w[i] * rho0[f] + (double)to_f(A[shift_elem(f, i, s, sp)]);
  }
};

// ---- layout conversion kernels ---------------------------------------------
// AoS staging [s][q] (physical sites s0.. s0+cnt) <-> SoA component f.
// Shifted storage keeps d_i = f_i - w_i rho0 rounded to the storage type.
template <class St>
__global__ void k_aos_to_soa(const double *__restrict__ stg, St *__restrict__ A,
                             St *__restrict__ B, int q, int f, int64_t sp, int s0, int cnt,
                             double rho0) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= cnt) return;
  for (int i = 0; i < q; ++i) {
    const St sv = to_storage<St>(stg[(int64_t)t * q + i], i, rho0);
    const int64_t o = pop_elem<St>(f, i, s0 + t, q, sp);
    A[o] = sv;
    B[o] = sv;
  }
}

template <class Ld>
__global__ void k_soa_to_aos(Ld ld, double *__restrict__ stg, int q, int f, int s0, int cnt) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= cnt) return;
  for (int i = 0; i < q; ++i) stg[(int64_t)t * q + i] = ld(f, i, s0 + t);
}

// D default-initialises a double to the quiet NaN 0x7FFC000000000000, and a
// component initialised with FluidInit.None keeps it (src/dlbc/lb/init.d:
// 263-270; the lb/init-2d reference data hold it).  Every population buffer
// starts out so: FP32 0x7FE00000 and FP16 0x7F00 widen to the same double.
template <class St>
__device__ inline St d_double_init();
template <>
__device__ inline double d_double_init<double>() { return __longlong_as_double(0x7FFC000000000000ll); }
template <>
__device__ inline float d_double_init<float>() { return __int_as_float(0x7FE00000); }
template <>
__device__ inline __half d_double_init<__half>() { return __ushort_as_half((unsigned short)0x7F00); }

template <class St>
__global__ void k_fill_default(St *__restrict__ A, St *__restrict__ B, int64_t n) {
  const St v = d_double_init<St>();
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    A[i] = v;
    B[i] = v;
  }
}

// A solid site's populations stay as they are, as in the reference
// (advection.d:84-86): the step kernels either skip it or (the split collide
// passes, where its cache line also holds fluid sites) store its own current
// populations again.  Skipping is right only while both buffers hold the same
// values there, so after a mask change mid-run the current buffer A is copied
// into B at every solid site.
template <class St>
__global__ void k_copy_solid(const St *__restrict__ A, St *__restrict__ B,
                             const uint8_t *__restrict__ mask, Geo g, int C, int q) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= g.sites || !mask[s]) return;
  for (int f = 0; f < C; ++f)
    for (int i = 0; i < q; ++i) {
      const int64_t o = pop_elem<St>(f, i, s, q, g.sp);
      B[o] = A[o];
    }
}

// Solid-neighbour word of every site: bit i = solid(x + c_i), bit 31 =
// solid(x) (periodic neighbours).
template <class L>
__global__ void k_solid_words(const uint8_t *__restrict__ mask, uint32_t *__restrict__ nbw,
                              Geo g) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= g.sites) return;
  int x, y, z;
  coords(s, g, x, y, z);
  uint32_t w = (uint32_t)(mask[s] != 0) << 31;
#pragma unroll
  for (int i = 0; i < L::q; ++i)
    w |= (uint32_t)(mask[site_of(wrap(x + L::c[i][0], g.nx), wrap(y + L::c[i][1], g.ny),
                                 wrap(z + L::c[i][2], g.nz), g)] != 0)
         << i;
  nbw[s] = w;
}

template <class L, class Ld>
__global__ void k_solid_rho(Ld ld, const uint8_t *__restrict__ mask, double *__restrict__ S,
                            float *__restrict__ P, int C, Geo g) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= g.sites || !mask[s]) return;
  for (int f = 0; f < C; ++f) {
    double n[L::q];
#pragma unroll
    for (int i = 0; i < L::q; ++i) n[i] = ld(f, i, s);
    const double rho = density<L>(n);
    S[f * g.sp + s] = rho;
    if (P) P[f * g.sp + s] = (float)rho;
  }
}

// Output fields (src/dlbc/lb/io.d:56-101).  kind 0: density of f1 (all sites);
// 1: raw velocity of f1 (lb/velocity.d:37-48, 106-117; 0 on solid);
// 2: colour rho_f1 - rho_f2 (lb/density.d:284-295; 0 on solid).
template <class L, class Ld>
__global__ void k_field(Ld ld, const uint8_t *__restrict__ mask, double *__restrict__ out, Geo g,
                        int kind, int f1, int f2, int has_solid) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= g.sites) return;
  double n[L::q];
#pragma unroll
  for (int i = 0; i < L::q; ++i) n[i] = ld(f1, i, s);
  const bool solid = has_solid && mask[s];
  if (kind == 0) {
    out[s] = density<L>(n);
  } else if (kind == 1) {
    double v[L::d];
#pragma unroll
    for (int j = 0; j < L::d; ++j) v[j] = 0.0;
    if (!solid) {
      const double rho = density<L>(n);
#pragma unroll
      for (int i = 0; i < L::q; ++i)
#pragma unroll
        for (int j = 0; j < L::d; ++j) v[j] += n[i] * static_cast<double>(L::c[i][j]);
#pragma unroll
      for (int j = 0; j < L::d; ++j) v[j] /= rho;
    }
#pragma unroll
    for (int j = 0; j < L::d; ++j) out[(int64_t)s * L::d + j] = v[j];
  } else {
    if (solid) {
      out[s] = 0.0;
    } else {
      double m[L::q];
#pragma unroll
      for (int i = 0; i < L::q; ++i) m[i] = ld(f2, i, s);
      out[s] = density<L>(n) - density<L>(m);
    }
  }
}

// Densities of both components and the fluid flag of the sites of a box
// [x0, x0+ex) x [y0, y0+ey) x [z0, z0+ez), row-major:
// out[k * (C+1) + f] = rho_f, out[k * (C+1) + C] = 1 (fluid) or 0 (solid).
template <class L, class Ld>
__global__ void k_box_density(Ld ld, const uint8_t *__restrict__ mask, double *__restrict__ out,
                              Geo g, int x0, int y0, int z0, int ex, int ey, int ez, int C,
                              int has_solid) {
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (t >= ex * ey * ez) return;
  const int s = site_of(x0 + t / (ey * ez), y0 + (t / ez) % ey, z0 + t % ez, g);
  for (int f = 0; f < C; ++f) {
    double n[L::q];
#pragma unroll
    for (int i = 0; i < L::q; ++i) n[i] = ld(f, i, s);
    out[(int64_t)t * (C + 1) + f] = density<L>(n);
  }
  out[(int64_t)t * (C + 1) + C] = (has_solid && mask[s]) ? 0.0 : 1.0;
}

// Deterministic double-double reduction over fluid sites.  Stage 1: every
// block sums a fixed contiguous chunk (thread-sequential, then across each
// wavefront by butterfly shuffles, then the block's four wave sums in order);
// stage 2: one block sums the block partials the same way.  what: 0 mass,
// 1..3 momentum.  The error-free additions are symmetric, so every lane of a
// butterfly ends with the same sum and the result does not depend on timing.
__device__ inline DD dd_plus(DD a, DD b) { return dd_add(a, b); }

__device__ inline DD wave_dd_sum(DD v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
    v = dd_plus(v, DD{__shfl_xor(v.hi, m, 64), __shfl_xor(v.lo, m, 64)});
  return v;
}

// Sum of the 256 threads' values: lane sums per wave, then waves 0..3 in order.
__device__ inline DD block_dd_sum(DD v) {
  __shared__ double wh[4], wl[4];
  v = wave_dd_sum(v);
  const int t = threadIdx.x;
  if ((t & 63) == 0) {
    wh[t >> 6] = v.hi;
    wl[t >> 6] = v.lo;
  }
  __syncthreads();
  DD r{wh[0], wl[0]};
  for (int w = 1; w < 4; ++w) r = dd_plus(r, DD{wh[w], wl[w]});
  return r;
}

template <class L, class Ld>
__global__ void __launch_bounds__(256) k_reduce1(Ld ld, const uint8_t *__restrict__ mask,
                                                 double *__restrict__ part, Geo g, int f,
                                                 int what, int chunk, int has_solid) {
  const int b = blockIdx.x, t = threadIdx.x;
  const int begin = b * chunk;
  const int end = min(begin + chunk, g.sites);
  DD acc{0.0, 0.0};
  for (int s = begin + t; s < end; s += 256) {
    if (has_solid && mask[s]) continue;
    double n[L::q];
#pragma unroll
    for (int i = 0; i < L::q; ++i) n[i] = ld(f, i, s);
    double v;
    if (what == 0) {
      v = density<L>(n);
    } else {
      v = 0.0;  // momentum (lb/momentum.d:37-52): sum n_i c_i in i order
#pragma unroll
      for (int i = 0; i < L::q; ++i) v += n[i] * static_cast<double>(L::c[i][what - 1]);
    }
    acc = dd_plus(acc, DD{v, 0.0});
  }
  const DD r = block_dd_sum(acc);
  if (t == 0) {
    part[2 * b] = r.hi;
    part[2 * b + 1] = r.lo;
  }
}

__global__ void __launch_bounds__(256) k_reduce2(const double *__restrict__ part, int nparts,
                                                 double *__restrict__ out) {
  const int t = threadIdx.x;
  DD acc{0.0, 0.0};
  for (int i = t; i < nparts; i += 256) acc = dd_plus(acc, DD{part[2 * i], part[2 * i + 1]});
  const DD r = block_dd_sum(acc);
  if (t == 0) {
    out[0] = r.hi;
    out[1] = r.lo;
  }
}

__global__ void k_eval_psi(const double *__restrict__ x, double *__restrict__ y, int64_t n,
                           int form) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = psi_of(form, x[i]);
}

// ---- dispatch over connectivity and component count ------------------------
template <template <class, int> class Fn, class... Args>
int dispatch(dlbc_gpu_ctx *c, Args... args) {
  auto go = [&](auto lat) -> int {
    using L = decltype(lat);
    switch (c->C) {
      case 1: return Fn<L, 1>::run(c, args...);
      case 2: return Fn<L, 2>::run(c, args...);
      case 3: return Fn<L, 3>::run(c, args...);
      case 4: return Fn<L, 4>::run(c, args...);
    }
    return fail(c, DLBC_ERR_UNSUPPORTED, "components %d", c->C);
  };
  if (c->d == 3) return go(D3Q19{});
  if (c->d == 2) return go(D2Q9{});
  if (c->q == 5) return go(D1Q5{});
  return go(D1Q3{});
}

template <class L>
int with_lattice(dlbc_gpu_ctx *c, int (*fn)(dlbc_gpu_ctx *, L)) {
  return fn(c, L{});
}

// Launch wrapper that optionally brackets the launch with events.
struct Timer {
  dlbc_gpu_ctx *c;
  bool on;
  int slot = 0;
  void before() {
    if (on) (void)hipEventRecord(c->ev[2 * slot], c->stream);
  }
  void after() {
    if (on) (void)hipEventRecord(c->ev[2 * slot + 1], c->stream);
    ++slot;
  }
};

// One step of a shifted-storage lattice: the tiled kernel of a batch over
// this lattice's own work list (8 x 32 tiles, x cut into chunks until the
// launch holds at least target_blocks workgroups, 512 by default: FP32
// arithmetic leaves room for two 512-thread workgroups per CU).
// The work list and the device table of that launch, brought up to date
// (copies only when the tile target, a parameter or a pointer changed).
int shifted_self_ready(dlbc_gpu_ctx *c) {
  const int target = c->target_blocks > 0 ? c->target_blocks : 512;
  if (c->items_target != target) {
    const int tiles = ((c->geo.ny + 7) / 8) * ((c->geo.nz + 31) / 32);
    const int want = std::max(1, std::min((target + tiles - 1) / tiles, c->geo.nx / 4));
    const int chunk = (c->geo.nx + want - 1) / want;
    std::vector<BatchItem> items;
    batch_items(&c->geo, 1, 8, chunk, items);
    if (c->d_items) HIPCHK(c, hipFree(c->d_items));
    c->d_items = nullptr;
    HIPCHK(c, hipMalloc(&c->d_items, sizeof(BatchItem) * items.size()));
    HIPCHK(c, hipMemcpyAsync(c->d_items, items.data(), sizeof(BatchItem) * items.size(),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->nitems = (int)items.size();
    c->items_target = target;
  }
  BatchLattice t;
  std::memset(&t, 0, sizeof t);
  t.pop[0] = c->bufX;
  t.pop[1] = c->bufY;
  t.nbw = c->nbw;
  t.S = c->S;
  t.P = c->P;
  t.g = c->geo;
  t.p = c->sparams;
  if (!c->self_valid || std::memcmp(&t, &c->self_tab, sizeof t)) {
    HIPCHK(c, hipMemcpyAsync(c->d_self, &t, sizeof t, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->self_tab = t;
    c->self_valid = true;
  }
  return DLBC_OK;
}

int launch_shifted_self(dlbc_gpu_ctx *c) {
  const int rc0 = shifted_self_ready(c);
  if (rc0) return rc0;
  const int parity = c->popA == c->bufX ? 0 : 1;
  const int grid = (c->nitems + 7) / 8 * 8;
  const int rc = c->storage == DLBC_STORAGE_F32_SHIFT
                     ? launch_batch_d3q19<float>(c->d_self, c->d_items, c->nitems, parity, 8,
                                                 grid, true, c->stream)
                     : launch_batch_d3q19<__half>(c->d_self, c->d_items, c->nitems, parity, 8,
                                                  grid, true, c->stream);
  return rc;
}

template <class L, int C>
struct TwoPassStep {
  static int run(dlbc_gpu_ctx *c, Timer *tm, double *acc_ms) {
    const int nb = blocks_for(c->geo.sites, 256);
    const double *A = static_cast<const double *>(c->popA);
    double *B = static_cast<double *>(c->popB);
    if (c->p.sc) {
      tm->before();
      hipLaunchKernelGGL((k_psi<L, C>), dim3(nb), dim3(256), 0, c->stream, A, c->mask, c->S,
                         c->geo, c->p.psi_form, c->has_solid);
      tm->after();
    }
    tm->before();
    hipLaunchKernelGGL((k_collide<L, C>), dim3(nb), dim3(256), 0, c->stream, A, B, c->mask, c->S,
                       c->geo, c->p);
    tm->after();
    (void)acc_ms;
    std::swap(c->popA, c->popB);
    return DLBC_OK;
  }
};

template <class L, int C>
struct FusedStep {
  static int run(dlbc_gpu_ctx *c, Timer *tm, double *acc_ms) {
    (void)acc_ms;
    if constexpr (std::is_same<L, D3Q19>::value && C == 2) {
      tm->before();
      int rc;
      if (c->storage == DLBC_STORAGE_F64)
        rc = launch_fused_d3q19(static_cast<const double *>(c->popA),
                                static_cast<double *>(c->popB), c->nbw, c->S, c->geo, c->p,
                                c->target_blocks, c->stream);
      else
        rc = launch_shifted_self(c);
      tm->after();
      if (rc) return fail(c, DLBC_ERR_UNSUPPORTED, "fused kernel: lattice shape not supported");
      std::swap(c->popA, c->popB);
      return DLBC_OK;
    } else {
      return fail(c, DLBC_ERR_UNSUPPORTED, "fused schedule needs D3Q19 with 2 components");
    }
  }
};

// Single pass over whole z-rows (D3Q19, two components, FP64).
template <class L, int C>
struct MarchStep {
  static int run(dlbc_gpu_ctx *c, Timer *tm, double *acc_ms) {
    (void)acc_ms;
    if constexpr (std::is_same<L, D3Q19>::value && C == 2) {
      if (c->storage != DLBC_STORAGE_F64)
        return fail(c, DLBC_ERR_UNSUPPORTED, "march schedule needs FP64 storage");
      tm->before();
      const int rc = launch_march_d3q19(static_cast<const double *>(c->popA),
                                        static_cast<double *>(c->popB), c->nbw, c->S, c->geo,
                                        c->p, c->stream, c->march_rows, c->march_hold,
                                        c->target_blocks);
      tm->after();
      if (rc) return fail(c, DLBC_ERR_UNSUPPORTED, "march kernel: lattice shape not supported");
      std::swap(c->popA, c->popB);
      return DLBC_OK;
    } else {
      return fail(c, DLBC_ERR_UNSUPPORTED, "march schedule needs D3Q19 with 2 components");
    }
  }
};

// Lane-per-component two-pass schedule (D3Q19, two components, FP64).
template <class L, int C>
struct SplitStep {
  static int run(dlbc_gpu_ctx *c, Timer *tm, double *acc_ms) {
    (void)acc_ms;
    if constexpr (std::is_same<L, D3Q19>::value && C == 2) {
      auto pass = [&](int k) {
        if (c->storage == DLBC_STORAGE_F64)
          launch_split_d3q19(static_cast<const double *>(c->popA),
                             static_cast<double *>(c->popB), c->nbw, c->S, c->geo, c->p,
                             c->stream, k);
        else if (c->storage == DLBC_STORAGE_F32_SHIFT)
          launch_ssplit_d3q19<float>(c->popA, c->popB, c->nbw, c->P, c->geo, c->sparams,
                                     c->stream, k);
        else
          launch_ssplit_d3q19<__half>(c->popA, c->popB, c->nbw, c->P, c->geo, c->sparams,
                                      c->stream, k);
      };
      if (c->p.sc) {
        tm->before();
        pass(0);
        tm->after();
      }
      tm->before();
      pass(1);
      tm->after();
      std::swap(c->popA, c->popB);
      return DLBC_OK;
    } else {
      return fail(c, DLBC_ERR_UNSUPPORTED, "split schedule needs D3Q19 with 2 components");
    }
  }
};

template <class L, int C>
struct AdvectOnly {
  static int run(dlbc_gpu_ctx *c) {
    if (c->storage != DLBC_STORAGE_F64)
      return fail(c, DLBC_ERR_UNSUPPORTED, "dlbc_gpu_advect needs FP64 storage");
    hipLaunchKernelGGL((k_advect<L, C>), dim3(blocks_for(c->geo.sites, 256)), dim3(256), 0,
                       c->stream, static_cast<const double *>(c->popA),
                       static_cast<double *>(c->popB), c->mask, c->geo, c->has_solid);
    std::swap(c->popA, c->popB);
    return DLBC_OK;
  }
};

// Call fn(lattice, loader) with the loader of the context's storage.
template <class Fn>
int with_loader(dlbc_gpu_ctx *c, Fn fn) {
  if (c->storage == DLBC_STORAGE_F64) {
    const LoadF64 ld{static_cast<const double *>(c->popA), c->geo.sp, c->q};
    if (c->d == 3) return fn(D3Q19{}, ld);
    if (c->d == 2) return fn(D2Q9{}, ld);
    if (c->q == 5) return fn(D1Q5{}, ld);
    return fn(D1Q3{}, ld);
  }
  if (c->storage == DLBC_STORAGE_F32_SHIFT) {
    const LoadShift<float> ld{static_cast<const float *>(c->popA), c->geo.sp,
                              {c->rho0[0], c->rho0[1], c->rho0[2], c->rho0[3]}};
    return fn(D3Q19{}, ld);
  }
  const LoadShift<__half> ld{static_cast<const __half *>(c->popA), c->geo.sp,
                             {c->rho0[0], c->rho0[1], c->rho0[2], c->rho0[3]}};
  return fn(D3Q19{}, ld);
}

int effective_schedule(const dlbc_gpu_ctx *c) {
  if (c->schedule != DLBC_SCHED_AUTO) return c->schedule;
  // shifted storage: the split pair (FP16 1.80 vs 1.94 ms on the 256^3 channel,
  // FP32 even); the tiled kernel (FUSED) remains selectable
  if (c->storage != DLBC_STORAGE_F64) return DLBC_SCHED_SPLIT;
  // D3Q19 two-component FP64: the lane-per-component two-pass kernels are the
  // fastest measured schedule (256^3 channel: 3.45 ms/step vs 4.51 ms fused)
  if (c->d == 3 && c->C == 2) return DLBC_SCHED_SPLIT;
  return DLBC_SCHED_TWO_PASS;
}

int prepare(dlbc_gpu_ctx *c) {
  if (!c->params_set) return fail(c, DLBC_ERR_STATE, "dlbc_gpu_set_params not called");
  for (int f = 0; f < c->C; ++f)
    if (!c->pops_uploaded[f]) return fail(c, DLBC_ERR_STATE, "component %d not uploaded", f);
  if (c->solid_rho_dirty) {
    if (c->has_solid) {
      const int nb = blocks_for(c->geo.sites, 256);
      with_loader(c, [&](auto lat, auto ld) {
        hipLaunchKernelGGL((k_solid_rho<decltype(lat), decltype(ld)>), dim3(nb), dim3(256), 0,
                           c->stream, ld, c->mask, c->S, c->P, c->C, c->geo);
        return 0;
      });
      HIPCHK(c, hipGetLastError());
    }
    c->solid_rho_dirty = false;
  }
  return DLBC_OK;
}

// One step under schedule `sched`: kernel launches on the library stream and
// the swap of the two population buffers, nothing else (so it can be captured).
int step_once(dlbc_gpu_ctx *c, int sched, Timer *tm) {
  tm->slot = 0;
  if (sched == DLBC_SCHED_FUSED) return dispatch<FusedStep>(c, tm, (double *)nullptr);
  if (sched == DLBC_SCHED_SPLIT) return dispatch<SplitStep>(c, tm, (double *)nullptr);
  if (sched == DLBC_SCHED_MARCH) return dispatch<MarchStep>(c, tm, (double *)nullptr);
  return dispatch<TwoPassStep>(c, tm, (double *)nullptr);
}

int do_steps(dlbc_gpu_ctx *c, int nsteps, Timer *tm) {
  int rc = prepare(c);
  if (rc) return rc;
  const int sched = effective_schedule(c);
  for (int t = 0; t < nsteps; ++t) {
    rc = step_once(c, sched, tm);
    if (rc) return rc;
  }
  HIPCHK(c, hipGetLastError());
  return DLBC_OK;
}

// Replays may still be in flight on the stream: wait for them before the
// executable graphs go.
void drop_graphs(dlbc_gpu_ctx *c) {
  if (!c->graph[0] && !c->graph[1]) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &g : c->graph) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
}

// Everything on the host that the kernel arguments of a step are taken from:
// schedule, launch shape, geometry (with the block order), parameters and the
// device table and work list of the shifted tiled kernel.  Buffers never move
// after creation; the two population buffers swap every step, which the
// parity index of the graphs covers.
std::vector<unsigned char> graph_key(const dlbc_gpu_ctx *c, int sched) {
  std::vector<unsigned char> k;
  auto put = [&](const void *p, size_t n) {
    const auto *b = static_cast<const unsigned char *>(p);
    k.insert(k.end(), b, b + n);
  };
  const int ints[6] = {sched,         c->storage,   c->has_solid,
                       c->target_blocks, c->march_rows, c->march_hold};
  put(ints, sizeof ints);
  put(&c->geo, sizeof c->geo);
  put(&c->p, sizeof c->p);
  put(&c->sparams, sizeof c->sparams);
  put(&c->d_items, sizeof c->d_items);
  put(&c->nitems, sizeof c->nitems);
  return k;
}

// host index of physical site (x,y,z) in a halo-padded array of the lattice
inline int64_t host_index(const dlbc_gpu_ctx *c, int halo, int x, int y, int z) {
  const int64_t hy = c->d >= 2 ? c->n[1] + 2 * halo : 1;
  const int64_t hz = c->d >= 3 ? c->n[2] + 2 * halo : 1;
  const int ox = halo, oy = c->d >= 2 ? halo : 0, oz = c->d >= 3 ? halo : 0;
  return ((int64_t)(x + ox) * hy + (y + oy)) * hz + (z + oz);
}

// Copy a per-site field (k values per site) from device `dev` (site-major)
// into a halo-padded host array.
int download_sitefield(dlbc_gpu_ctx *c, const double *dev, int k, double *host, int halo) {
  std::vector<double> buf((size_t)c->geo.sites * k);
  HIPCHK(c, hipMemcpyAsync(buf.data(), dev, buf.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  size_t o = 0;
  for (int x = 0; x < c->n[0]; ++x)
    for (int y = 0; y < c->n[1]; ++y)
      for (int z = 0; z < c->n[2]; ++z) {
        std::memcpy(host + host_index(c, halo, x, y, z) * k, buf.data() + o * k,
                    sizeof(double) * k);
        ++o;
      }
  return DLBC_OK;
}

int field_launch(dlbc_gpu_ctx *c, double *out, int kind, int f1, int f2) {
  const int nb = blocks_for(c->geo.sites, 256);
  with_loader(c, [&](auto lat, auto ld) {
    hipLaunchKernelGGL((k_field<decltype(lat), decltype(ld)>), dim3(nb), dim3(256), 0, c->stream,
                       ld, c->mask, out, c->geo, kind, f1, f2, c->has_solid);
    return 0;
  });
  HIPCHK(c, hipGetLastError());
  return DLBC_OK;
}

// Sites per block of the first reduction stage, and the doubles of device
// memory a reduction needs: two per block partial and two for the result.
constexpr int kReduceChunk = 1 << 16;
inline size_t reduce_doubles(const dlbc_gpu_ctx *c) {
  return 2 * (size_t)blocks_for(c->geo.sites, kReduceChunk) + 2;
}

// The two reduction stages into device memory `work` (reduce_doubles(c)
// doubles), the result copied to out2 on the host.
int reduce_into(dlbc_gpu_ctx *c, int f, int what, double *work, double *out2) {
  const int chunk = kReduceChunk;
  const int nparts = blocks_for(c->geo.sites, chunk);
  double *part = work, *res = work + 2 * nparts;
  with_loader(c, [&](auto lat, auto ld) {
    hipLaunchKernelGGL((k_reduce1<decltype(lat), decltype(ld)>), dim3(nparts), dim3(256), 0,
                       c->stream, ld, c->mask, part, c->geo, f, what, chunk, c->has_solid);
    return 0;
  });
  hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(256), 0, c->stream, part, nparts, res);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(out2, res, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DLBC_OK;
}

int reduce_launch(dlbc_gpu_ctx *c, int f, int what, double *out2) {
  int rc = ensure_scratch(c, sizeof(double) * reduce_doubles(c));
  if (rc) return rc;
  return reduce_into(c, f, what, c->scratch, out2);
}

// ---- the allocation-free mass reduction --------------------------------------
// dlbc_gpu_total_mass_noalloc keeps its block partials in this file-scope
// array instead of context scratch.  Every GPU holds its own copy, at its own
// address, until the process ends; every call on a GPU shares it, so calls
// that overlap on one GPU race.  In host code the array's name is a host-side
// stand-in, never a device address: kernels get the address that
// hipGetSymbolAddress returned for the device (noalloc_ptr).  Nothing in it
// must start at zero: every block writes its own partial, the second stage
// writes the result.
constexpr int kNoAllocDoubles = 8192;  // 64 KiB: 4095 partials, 2^16 sites each
static __device__ __align__(256) double noalloc_scratch[kNoAllocDoubles];

// The largest lattice it serves: (8192 - 2) / 2 partials of kReduceChunk sites.
constexpr int64_t kNoAllocMaxSites = (int64_t)((kNoAllocDoubles - 2) / 2) * kReduceChunk;

constexpr int kNoAllocMaxDevices = 64;
std::mutex noalloc_mu;
double *noalloc_ptr[kNoAllocMaxDevices] = {};

// Look up noalloc_scratch's device address on c->device, once per device; the
// caller has made that device current (dlbc_gpu_create, after hipSetDevice).
int noalloc_setup(dlbc_gpu_ctx *c) {
  if (c->device < 0 || c->device >= kNoAllocMaxDevices)
    return fail(c, DLBC_ERR_UNSUPPORTED, "device %d: the reduction scratch covers devices 0..%d",
                c->device, kNoAllocMaxDevices - 1);
  std::lock_guard<std::mutex> lock(noalloc_mu);
  if (noalloc_ptr[c->device]) return DLBC_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
  if (e != hipSuccess || !p)
    return fail(c, DLBC_ERR_HIP, "hipGetSymbolAddress(noalloc_scratch) failed on device %d: %s",
                c->device, hipGetErrorString(e));
  noalloc_ptr[c->device] = static_cast<double *>(p);
  return DLBC_OK;
}

// volumeAveragedDensity (src/dlbc/lb/laplace.d:236-281) with volumeAverage 1:
// mean density per component over the fluid sites of the 3^d cube centred at
// `centre` (clipped at the lattice edge, as the reference's cube does not
// wrap), summed in row-major order.
int cube_density(dlbc_gpu_ctx *c, const int *centre, double *avg) {
  int lo[3] = {0, 0, 0}, ext[3] = {1, 1, 1};
  for (int vd = 0; vd < c->d; ++vd) {
    lo[vd] = std::max(0, centre[vd] - 1);
    ext[vd] = std::min(c->n[vd] - 1, centre[vd] + 1) - lo[vd] + 1;
  }
  const int cnt = ext[0] * ext[1] * ext[2];
  const int C = c->C;
  int rc = ensure_scratch(c, sizeof(double) * cnt * (C + 1));
  if (rc) return rc;
  with_loader(c, [&](auto lat, auto ld) {
    hipLaunchKernelGGL((k_box_density<decltype(lat), decltype(ld)>), dim3(1), dim3(64), 0,
                       c->stream, ld, c->mask, c->scratch, c->geo, lo[0], lo[1], lo[2], ext[0],
                       ext[1], ext[2], C, c->has_solid);
    return 0;
  });
  HIPCHK(c, hipGetLastError());
  std::vector<double> buf((size_t)cnt * (C + 1));
  HIPCHK(c, hipMemcpyAsync(buf.data(), c->scratch, buf.size() * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int f = 0; f < C; ++f) {
    double sum = 0.0;
    int n = 0;
    for (int k = 0; k < cnt; ++k)
      if (buf[(size_t)k * (C + 1) + C] != 0.0) {
        sum += buf[(size_t)k * (C + 1) + f];
        ++n;
      }
    avg[f] = sum / n;
  }
  return DLBC_OK;
}

// pressurePsi (src/dlbc/lb/density.d:316-326).
double sc_pressure(const dlbc_gpu_ctx *c, const double *den) {
  auto psi = [&](double r) { return c->p.psi_form == 0 ? r : psi_exponential(r); };
  double pr = 0.0;
  for (int i = 0; i < c->C; ++i) {
    pr += den[i];
    for (int j = 0; j < c->C; ++j) pr += 0.5 * c->p.gcc[i][j] * psi(den[i]) * psi(den[j]);
  }
  return pr * kCss;
}

// Droplet deformation of a 2-D lattice (src/dlbc/lb/laplace.d:118-199):
// inertia tensor of the sites with rho_0 > 0.35 about their centre of mass.
double deformation_2d(const dlbc_gpu_ctx *c, const std::vector<double> &rho,
                      const std::vector<uint8_t> &solid) {
  const double cutoff = 0.35;
  const int nx = c->n[0], ny = c->n[1];
  double rx = 0.0, ry = 0.0, mass = 0.0;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y) {
      const double den = rho[(size_t)x * ny + y];
      if (!solid[(size_t)x * ny + y] && den > cutoff) {
        mass += den;
        rx += den * x;
        ry += den * y;
      }
    }
  rx /= mass;
  ry /= mass;
  double ixx = 0.0, ixy = 0.0, iyy = 0.0;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y) {
      const double den = rho[(size_t)x * ny + y];
      if (!solid[(size_t)x * ny + y] && den > cutoff) {
        const double xx = x - rx, yy = y - ry;
        ixx += den * (yy * yy);
        ixy += den * (xx * yy);
        iyy += den * (xx * xx);
      }
    }
  const double tr = ixx + iyy, det = ixx * iyy - ixy * ixy;
  const double ev1 = 0.5 * tr + std::sqrt(0.25 * tr * tr - det);
  const double ev2 = 0.5 * tr - std::sqrt(0.25 * tr * tr - det);
  const double l = std::sqrt(5.0 * std::max(ev1, ev2) / mass);
  const double b = std::sqrt(5.0 * std::min(ev1, ev2) / mass);
  const double D = (l - b) / (l + b);
  return std::isnan(D) ? 0.0 : D;
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

int dlbc_gpu_create(dlbc_gpu_ctx **out, int d, int q, int components, const int64_t *n,
                    int storage, int device) {
  if (!out || !n) return DLBC_ERR_ARG;
  *out = nullptr;
  if (!((d == 3 && q == 19) || (d == 2 && q == 9) || (d == 1 && (q == 3 || q == 5))))
    return DLBC_ERR_UNSUPPORTED;
  if (components < 1 || components > kMaxC) return DLBC_ERR_UNSUPPORTED;
  if (storage < DLBC_STORAGE_F64 || storage > DLBC_STORAGE_F16_SHIFT) return DLBC_ERR_ARG;
  if (storage != DLBC_STORAGE_F64 && !(d == 3 && q == 19 && components == 2))
    return DLBC_ERR_UNSUPPORTED;  // shifted storage: the D3Q19 two-component kernel only
  int64_t sites = 1;
  // the periodic images the reference's halo holds: n >= 2 (halo 2), and
  // n >= 4 for D1Q5 (halo 4, src/dlbc/lattice.d:96-103)
  const int64_t nmin = q == 5 ? 4 : 2;
  for (int i = 0; i < d; ++i) {
    if (n[i] < nmin || n[i] > (1 << 20)) return DLBC_ERR_ARG;
    sites *= n[i];
  }
  if (sites > (int64_t)1 << 30) return DLBC_ERR_UNSUPPORTED;
  auto *c = new dlbc_gpu_ctx();
  c->d = d;
  c->q = q;
  c->C = components;
  c->storage = storage;
  c->device = device;
  for (int i = 0; i < d; ++i) c->n[i] = (int)n[i];
  c->geo.nx = c->n[0];
  c->geo.ny = c->n[1];
  c->geo.nz = c->n[2];
  c->geo.sites = (int)sites;
  c->geo.sp = (sites + 63) / 64 * 64;
  geo_set_divisors(c->geo);
  apply_block_order(c, -1);
  auto bail = [&](int code) {
    dlbc_gpu_destroy(c);
    return code;
  };
  if (storage != DLBC_STORAGE_F64 && (c->n[0] < 4 || c->n[1] < 4 || c->n[2] < 4))
    return bail(DLBC_ERR_UNSUPPORTED);  // the tiled kernel needs n >= 4 on every axis
  c->elem = storage == DLBC_STORAGE_F64 ? 8 : (storage == DLBC_STORAGE_F32_SHIFT ? 4 : 2);
  if (hipSetDevice(device) != hipSuccess) return bail(DLBC_ERR_HIP);
  if (const int rc = noalloc_setup(c)) return bail(rc);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(DLBC_ERR_HIP);
  const size_t popb = (size_t)c->elem * components * q * c->geo.sp;
  if (hipMalloc(&c->popA, popb) != hipSuccess) return bail(DLBC_ERR_NOMEM);
  if (hipMalloc(&c->popB, popb) != hipSuccess) return bail(DLBC_ERR_NOMEM);
  c->bufX = c->popA;
  c->bufY = c->popB;
  if (storage != DLBC_STORAGE_F64 && hipMalloc(&c->d_self, sizeof(BatchLattice)) != hipSuccess)
    return bail(DLBC_ERR_NOMEM);
  if (storage != DLBC_STORAGE_F64 &&
      hipMalloc(&c->P, sizeof(float) * components * c->geo.sp) != hipSuccess)
    return bail(DLBC_ERR_NOMEM);
  if (hipMalloc(&c->S, sizeof(double) * components * c->geo.sp) != hipSuccess)
    return bail(DLBC_ERR_NOMEM);
  if (hipMalloc(&c->mask, c->geo.sp) != hipSuccess) return bail(DLBC_ERR_NOMEM);
  if (hipMalloc(&c->nbw, sizeof(uint32_t) * c->geo.sp) != hipSuccess) return bail(DLBC_ERR_NOMEM);
  c->dev_bytes = 2 * popb + sizeof(double) * components * c->geo.sp + 5 * c->geo.sp +
                 (c->P ? sizeof(float) * components * c->geo.sp : 0);
  if (hipMemsetAsync(c->mask, 0, c->geo.sp, c->stream) != hipSuccess) return bail(DLBC_ERR_HIP);
  if (hipMemsetAsync(c->nbw, 0, sizeof(uint32_t) * c->geo.sp, c->stream) != hipSuccess)
    return bail(DLBC_ERR_HIP);
  if (hipMemsetAsync(c->S, 0, sizeof(double) * components * c->geo.sp, c->stream) != hipSuccess)
    return bail(DLBC_ERR_HIP);
  {
    const int64_t nel = (int64_t)components * q * c->geo.sp;
    const dim3 grid((unsigned)std::min<int64_t>(blocks_for(nel, 256), 8192)), block(256);
    if (storage == DLBC_STORAGE_F64)
      hipLaunchKernelGGL(k_fill_default<double>, grid, block, 0, c->stream,
                         static_cast<double *>(c->popA), static_cast<double *>(c->popB), nel);
    else if (storage == DLBC_STORAGE_F32_SHIFT)
      hipLaunchKernelGGL(k_fill_default<float>, grid, block, 0, c->stream,
                         static_cast<float *>(c->popA), static_cast<float *>(c->popB), nel);
    else
      hipLaunchKernelGGL(k_fill_default<__half>, grid, block, 0, c->stream,
                         static_cast<__half *>(c->popA), static_cast<__half *>(c->popB), nel);
    if (hipGetLastError() != hipSuccess) return bail(DLBC_ERR_HIP);
  }
  for (auto &e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) return bail(DLBC_ERR_HIP);
  c->events_ready = true;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(DLBC_ERR_HIP);
  *out = c;
  return DLBC_OK;
}

int dlbc_gpu_set_params(dlbc_gpu_ctx *c, const double *tau, const double *gcc, const double *gwc,
                        const double *global_acc, const double *body_force, int eq_dist_form,
                        int psi_form, int enable_shan_chen, const double *rho0) {
  if (!c) return DLBC_ERR_ARG;
  if (!tau) return fail(c, DLBC_ERR_ARG, "tau is required");
  if (eq_dist_form < 0 || eq_dist_form > 2 || psi_form < 0 || psi_form > 1)
    return fail(c, DLBC_ERR_ARG, "bad eq_dist_form/psi_form");
  if (enable_shan_chen && !gcc) return fail(c, DLBC_ERR_ARG, "gcc required with Shan-Chen");
  (void)rho0;
  StepParams p{};
  for (int f = 0; f < c->C; ++f) {
    p.tau[f] = tau[f];
    p.omega[f] = 1.0 / tau[f];  // collision.d:199
    p.one_minus_omega[f] = 1.0 - p.omega[f];
    p.gwc[f] = gwc ? gwc[f] : 0.0;
    for (int g = 0; g < c->C; ++g) p.gcc[f][g] = gcc ? gcc[f * c->C + g] : 0.0;
    for (int k = 0; k < c->d; ++k) {
      p.body[f][k] = body_force ? body_force[f * c->d + k] : 0.0;
      if (p.body[f][k] != 0.0) p.has_body = 1;
    }
  }
  for (int k = 0; k < c->d; ++k) p.gacc[k] = global_acc ? global_acc[k] : 0.0;
  p.eq_form = eq_dist_form;
  p.psi_form = psi_form;
  p.sc = enable_shan_chen ? 1 : 0;
  p.has_solid = c->has_solid;
  // Shifted storage keeps f_i - w_i rho0_f: once a component's populations are
  // on the device its rho0 is part of their encoding and cannot change (NULL
  // keeps the current values, 1.0 until first set).
  if (rho0)
    for (int f = 0; f < c->C; ++f)
      if (c->storage != DLBC_STORAGE_F64 && c->pops_uploaded[f] && rho0[f] != c->rho0[f])
        return fail(c, DLBC_ERR_STATE,
                    "rho0 of component %d cannot change once its populations are on the "
                    "device (%g stored, %g given)", f, c->rho0[f], rho0[f]);
  c->p = p;
  if (rho0)
    for (int f = 0; f < c->C; ++f) c->rho0[f] = rho0[f];
  if (c->storage != DLBC_STORAGE_F64) {
    ShiftParams sp{};
    for (int f = 0; f < 2; ++f) {
      sp.tau[f] = (float)p.tau[f];
      sp.itau[f] = 1.0f / sp.tau[f];
      sp.omega[f] = (float)p.omega[f];
      sp.rho0[f] = (float)c->rho0[f];
      sp.gwc[f] = (float)p.gwc[f];
      for (int g = 0; g < 2; ++g) sp.gcc[f][g] = (float)p.gcc[f][g];
      for (int k = 0; k < 3; ++k) sp.body[f][k] = (float)p.body[f][k];
    }
    for (int k = 0; k < 3; ++k) sp.gacc[k] = (float)p.gacc[k];
    sp.eq_form = p.eq_form;
    sp.psi_form = p.psi_form;
    sp.sc = p.sc;
    sp.has_solid = p.has_solid;
    sp.has_body = p.has_body;
    c->sparams = sp;
  }
  c->params_set = true;
  return DLBC_OK;
}

static int mask_ready(dlbc_gpu_ctx *c, int solid);

int dlbc_gpu_upload_mask(dlbc_gpu_ctx *c, const int32_t *host, int halo) {
  if (!c || !host || halo < 0) return DLBC_ERR_ARG;
  std::vector<uint8_t> m((size_t)c->geo.sites);
  int solid = 0;
  size_t o = 0;
  for (int x = 0; x < c->n[0]; ++x)
    for (int y = 0; y < c->n[1]; ++y)
      for (int z = 0; z < c->n[2]; ++z) {
        const int32_t v = host[host_index(c, halo, x, y, z)];
        if (v != 0 && v != 1) return fail(c, DLBC_ERR_ARG, "mask value %d is not 0 or 1", v);
        m[o++] = (uint8_t)v;
        solid |= v;
      }
  HIPCHK(c, hipMemcpyAsync(c->mask, m.data(), m.size(), hipMemcpyHostToDevice, c->stream));
  return mask_ready(c, solid);
}

// The mask bytes are on the device: derive the solid-neighbour words and the
// solid flag every step kernel reads.
static int mask_ready(dlbc_gpu_ctx *c, int solid) {
  {
    const int nb = blocks_for(c->geo.sites, 256);
    if (c->d == 3)
      hipLaunchKernelGGL(k_solid_words<D3Q19>, dim3(nb), dim3(256), 0, c->stream, c->mask, c->nbw,
                         c->geo);
    else if (c->d == 2)
      hipLaunchKernelGGL(k_solid_words<D2Q9>, dim3(nb), dim3(256), 0, c->stream, c->mask, c->nbw,
                         c->geo);
    else if (c->q == 5)
      hipLaunchKernelGGL(k_solid_words<D1Q5>, dim3(nb), dim3(256), 0, c->stream, c->mask, c->nbw,
                         c->geo);
    else
      hipLaunchKernelGGL(k_solid_words<D1Q3>, dim3(nb), dim3(256), 0, c->stream, c->mask, c->nbw,
                         c->geo);
    HIPCHK(c, hipGetLastError());
  }
  if (solid) {
    const int nb = blocks_for(c->geo.sites, 256);
    if (c->storage == DLBC_STORAGE_F64)
      hipLaunchKernelGGL(k_copy_solid<double>, dim3(nb), dim3(256), 0, c->stream,
                         static_cast<const double *>(c->popA), static_cast<double *>(c->popB),
                         c->mask, c->geo, c->C, c->q);
    else if (c->storage == DLBC_STORAGE_F32_SHIFT)
      hipLaunchKernelGGL(k_copy_solid<float>, dim3(nb), dim3(256), 0, c->stream,
                         static_cast<const float *>(c->popA), static_cast<float *>(c->popB),
                         c->mask, c->geo, c->C, c->q);
    else
      hipLaunchKernelGGL(k_copy_solid<__half>, dim3(nb), dim3(256), 0, c->stream,
                         static_cast<const __half *>(c->popA), static_cast<__half *>(c->popB),
                         c->mask, c->geo, c->C, c->q);
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->has_solid = solid;
  c->p.has_solid = solid;
  c->sparams.has_solid = solid;
  c->solid_rho_dirty = true;
  return DLBC_OK;
}

int dlbc_gpu_upload_populations(dlbc_gpu_ctx *c, int f, const double *host, int halo) {
  if (!c || !host || halo < 0 || f < 0 || f >= c->C) return DLBC_ERR_ARG;
  const int q = c->q;
  const int plane = c->n[1] * c->n[2];
  const int xchunk = std::max(1, std::min(c->n[0], (int)((64u << 20) / (sizeof(double) * q * plane))));
  int rc = ensure_scratch(c, sizeof(double) * (size_t)q * plane * xchunk);
  if (rc) return rc;
  std::vector<double> stg((size_t)q * plane * xchunk);
  for (int x0 = 0; x0 < c->n[0]; x0 += xchunk) {
    const int nxc = std::min(xchunk, c->n[0] - x0);
    size_t o = 0;
    for (int x = x0; x < x0 + nxc; ++x)
      for (int y = 0; y < c->n[1]; ++y) {
        // a z-row of physical sites is contiguous in the host array
        std::memcpy(stg.data() + o * q, host + host_index(c, halo, x, y, 0) * q,
                    sizeof(double) * q * c->n[2]);
        o += c->n[2];
      }
    HIPCHK(c, hipMemcpyAsync(c->scratch, stg.data(), sizeof(double) * o * q,
                             hipMemcpyHostToDevice, c->stream));
    const dim3 grid(blocks_for((int64_t)o, 256));
    if (c->storage == DLBC_STORAGE_F64)
      hipLaunchKernelGGL(k_aos_to_soa<double>, grid, dim3(256), 0, c->stream, c->scratch,
                         static_cast<double *>(c->popA), static_cast<double *>(c->popB), q, f,
                         c->geo.sp, x0 * plane, (int)o, 0.0);
    else if (c->storage == DLBC_STORAGE_F32_SHIFT)
      hipLaunchKernelGGL(k_aos_to_soa<float>, grid, dim3(256), 0, c->stream, c->scratch,
                         static_cast<float *>(c->popA), static_cast<float *>(c->popB), q, f,
                         c->geo.sp, x0 * plane, (int)o, c->rho0[f]);
    else
      hipLaunchKernelGGL(k_aos_to_soa<__half>, grid, dim3(256), 0, c->stream, c->scratch,
                         static_cast<__half *>(c->popA), static_cast<__half *>(c->popB), q, f,
                         c->geo.sp, x0 * plane, (int)o, c->rho0[f]);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  c->pops_uploaded[f] = true;
  c->solid_rho_dirty = true;
  return DLBC_OK;
}

// Launch k_init_fluid over sites s0 .. s0+cnt-1 in the lattice's storage.
static int launch_init_fluid(dlbc_gpu_ctx *c, int s0, int cnt, const double *vals,
                             const FluidInitArgs &a, const ShapeSpec &sh) {
  const dim3 grid(blocks_for(cnt, 256)), block(256);
  if (c->storage == DLBC_STORAGE_F64)
    hipLaunchKernelGGL(k_init_fluid<double>, grid, block, 0, c->stream,
                       static_cast<double *>(c->popA), static_cast<double *>(c->popB), c->geo, s0,
                       cnt, vals, a, sh);
  else if (c->storage == DLBC_STORAGE_F32_SHIFT)
    hipLaunchKernelGGL(k_init_fluid<float>, grid, block, 0, c->stream,
                       static_cast<float *>(c->popA), static_cast<float *>(c->popB), c->geo, s0,
                       cnt, vals, a, sh);
  else
    hipLaunchKernelGGL(k_init_fluid<__half>, grid, block, 0, c->stream,
                       static_cast<__half *>(c->popA), static_cast<__half *>(c->popB), c->geo, s0,
                       cnt, vals, a, sh);
  HIPCHK(c, hipGetLastError());
  return DLBC_OK;
}

int dlbc_gpu_init_fluid(dlbc_gpu_ctx *c, int f, int halo, int eq_dist_form, int kind,
                        const double *dens, int ndens, double perturb, dlbc_rng *rng,
                        double radius, const double *offset, const double *separation, int axis,
                        double thickness, const double *widths) {
  if (!c || f < 0 || f >= c->C || halo < 0) return DLBC_ERR_ARG;
  FluidInitArgs a{};
  a.q = c->q;
  a.f = f;
  a.halo = halo;
  a.rho0 = c->rho0[f];
  if (!eq_unity(c->d, c->q, eq_dist_form, a.u))
    return fail(c, DLBC_ERR_ARG, "eq_dist_form %d is not a known form", eq_dist_form);
  if (kind == DLBC_INIT_NONE || kind == DLBC_INIT_FILE) return DLBC_OK;  // nothing is written
  ShapeSpec sh{};
  const bool random = kind == DLBC_INIT_CONST_RANDOM || kind == DLBC_INIT_EQDIST_RANDOM ||
                      kind == DLBC_INIT_EQDIST_PERTURB || kind == DLBC_INIT_EQDIST_PERTURB_FRAC;
  if (kind == DLBC_INIT_CONST || kind == DLBC_INIT_EQDIST) {  // init.d:26-37, 52-58
    if (!dens || ndens < 1) return fail(c, DLBC_ERR_ARG, "fluid init needs a density");
    a.mode = kInitConst;
    a.value = dens[0];
    if (kind == DLBC_INIT_CONST)
      for (int i = 0; i < a.q; ++i) a.u[i] = 1.0;
  } else if (random) {  // init.d:39-50, 60-82
    if (!dens || ndens < 1 || !rng) return fail(c, DLBC_ERR_ARG, "random fluid init needs a "
                                                                 "density and a generator");
    a.mode = kind == DLBC_INIT_CONST_RANDOM ? kInitPop : kInitSite;
  } else if (is_shape_kind(kind)) {
    const int64_t n64[3] = {c->n[0], c->n[1], c->n[2]};
    if (!resolve_shape(sh, kind, c->d, n64, halo, dens, ndens, radius, offset, separation, axis,
                       thickness, widths))
      return fail(c, DLBC_ERR_ARG, "invalid shape arguments for fluid init kind %d", kind);
    a.mode = kInitShape;
  } else {
    return fail(c, DLBC_ERR_ARG, "unknown fluid init kind %d", kind);
  }
  if (!random) {
    int rc = launch_init_fluid(c, 0, c->geo.sites, nullptr, a, sh);
    if (rc) return rc;
  } else {
    // The reference draws over its halo-padded field in row-major order
    // (one draw per site, per population for ConstRandom); the halo's draws
    // are discarded.  Physical x-planes go to the device in chunks.
    const int per = a.mode == kInitPop ? a.q : 1;
    const int64_t h = halo;
    const int64_t nh[3] = {c->n[0] + 2 * h, c->d >= 2 ? c->n[1] + 2 * h : 1,
                           c->d >= 3 ? c->n[2] + 2 * h : 1};
    const int64_t oy = c->d >= 2 ? h : 0, oz = c->d >= 3 ? h : 0;
    const int plane = c->n[1] * c->n[2];
    const int xchunk =
        std::max(1, std::min(c->n[0], (int)((64u << 20) / (sizeof(double) * per * plane))));
    int rc = ensure_scratch(c, sizeof(double) * (size_t)per * plane * xchunk);
    if (rc) return rc;
    std::vector<double> stg((size_t)per * plane * xchunk);
    size_t o = 0;
    int x0 = 0;  // first physical plane in the staging buffer
    for (int64_t px = 0; px < nh[0]; ++px) {
      const bool xin = px >= h && px < h + c->n[0];
      for (int64_t py = 0; py < nh[1]; ++py)
        for (int64_t pz = 0; pz < nh[2]; ++pz) {
          const bool in = xin && py >= oy && py < oy + c->n[1] && pz >= oz && pz < oz + c->n[2];
          if (per == 1) {
            const double v = random_density(kind, dens[0], perturb, rng);
            if (in) stg[o++] = v;
          } else {
            for (int i = 0; i < per; ++i) {
              const double v = dens[0] * dlbc_rng_uniform(rng, 0.0, 2.0);
              if (in) stg[o++] = v;
            }
          }
        }
      const int xdone = (int)(px - h) + 1;  // physical planes staged so far
      if (xin && (xdone - x0 == xchunk || xdone == c->n[0])) {
        const int cnt = (xdone - x0) * plane;
        HIPCHK(c, hipMemcpyAsync(c->scratch, stg.data(), sizeof(double) * o,
                                 hipMemcpyHostToDevice, c->stream));
        rc = launch_init_fluid(c, x0 * plane, cnt, c->scratch, a, sh);
        if (rc) return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));  // the staging buffer is refilled next
        o = 0;
        x0 = xdone;
      }
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->pops_uploaded[f] = true;
  c->solid_rho_dirty = true;
  return DLBC_OK;
}

int dlbc_gpu_init_mask(dlbc_gpu_ctx *c, int kind, int axis, int wall_offset,
                       const int *dim_has_wall, int wall_thickness) {
  if (!c) return DLBC_ERR_ARG;
  const int64_t n64[3] = {c->n[0], c->n[1], c->n[2]};
  MaskSpec m;
  if (!resolve_mask(m, c->d, n64, 0, kind, axis, wall_offset, dim_has_wall, wall_thickness))
    return fail(c, DLBC_ERR_ARG, "invalid mask init (kind %d, axis %d)", kind, axis);
  hipLaunchKernelGGL(k_init_mask, dim3(blocks_for(c->geo.sites, 256)), dim3(256), 0, c->stream,
                     c->mask, c->geo, m);
  HIPCHK(c, hipGetLastError());
  std::vector<uint8_t> back((size_t)c->geo.sites);
  HIPCHK(c, hipMemcpyAsync(back.data(), c->mask, back.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int solid = 0;
  for (uint8_t v : back) solid |= v;
  return mask_ready(c, solid);
}

int dlbc_gpu_init_eq_dist_wall(dlbc_gpu_ctx *c, int f) {
  if (!c || f < 0 || f >= c->C) return DLBC_ERR_ARG;
  if (!c->has_solid) return DLBC_OK;
  const dim3 grid(blocks_for(c->geo.sites, 256)), block(256);
  if (c->storage == DLBC_STORAGE_F64)
    hipLaunchKernelGGL(k_eq_dist_wall<double>, grid, block, 0, c->stream,
                       static_cast<double *>(c->popA), static_cast<double *>(c->popB), c->mask,
                       c->geo, c->q, f, c->rho0[f]);
  else if (c->storage == DLBC_STORAGE_F32_SHIFT)
    hipLaunchKernelGGL(k_eq_dist_wall<float>, grid, block, 0, c->stream,
                       static_cast<float *>(c->popA), static_cast<float *>(c->popB), c->mask,
                       c->geo, c->q, f, c->rho0[f]);
  else
    hipLaunchKernelGGL(k_eq_dist_wall<__half>, grid, block, 0, c->stream,
                       static_cast<__half *>(c->popA), static_cast<__half *>(c->popB), c->mask,
                       c->geo, c->q, f, c->rho0[f]);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->solid_rho_dirty = true;
  return DLBC_OK;
}

int dlbc_gpu_download_populations(dlbc_gpu_ctx *c, int f, double *host, int halo) {
  if (!c || !host || halo < 0 || f < 0 || f >= c->C) return DLBC_ERR_ARG;
  const int q = c->q;
  const int plane = c->n[1] * c->n[2];
  const int xchunk = std::max(1, std::min(c->n[0], (int)((64u << 20) / (sizeof(double) * q * plane))));
  int rc = ensure_scratch(c, sizeof(double) * (size_t)q * plane * xchunk);
  if (rc) return rc;
  std::vector<double> stg((size_t)q * plane * xchunk);
  for (int x0 = 0; x0 < c->n[0]; x0 += xchunk) {
    const int nxc = std::min(xchunk, c->n[0] - x0);
    const int cnt = nxc * plane;
    with_loader(c, [&](auto, auto ld) {
      hipLaunchKernelGGL((k_soa_to_aos<decltype(ld)>), dim3(blocks_for(cnt, 256)), dim3(256), 0,
                         c->stream, ld, c->scratch, q, f, x0 * plane, cnt);
      return 0;
    });
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(stg.data(), c->scratch, sizeof(double) * cnt * q,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    size_t o = 0;
    for (int x = x0; x < x0 + nxc; ++x)
      for (int y = 0; y < c->n[1]; ++y) {
        std::memcpy(host + host_index(c, halo, x, y, 0) * q, stg.data() + o * q,
                    sizeof(double) * q * c->n[2]);
        o += c->n[2];
      }
  }
  return DLBC_OK;
}

int dlbc_gpu_download_mask(dlbc_gpu_ctx *c, int32_t *host, int halo) {
  if (!c || !host || halo < 0) return DLBC_ERR_ARG;
  std::vector<uint8_t> m((size_t)c->geo.sites);
  HIPCHK(c, hipMemcpyAsync(m.data(), c->mask, m.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  size_t o = 0;
  for (int x = 0; x < c->n[0]; ++x)
    for (int y = 0; y < c->n[1]; ++y)
      for (int z = 0; z < c->n[2]; ++z) host[host_index(c, halo, x, y, z)] = m[o++];
  return DLBC_OK;
}

int dlbc_gpu_step(dlbc_gpu_ctx *c, int nsteps) {
  if (!c || nsteps < 0) return DLBC_ERR_ARG;
  Timer tm{c, false};
  return do_steps(c, nsteps, &tm);
}

// Steps per captured graph (even).
constexpr int kGraphSteps = 32;

int dlbc_gpu_step_graph(dlbc_gpu_ctx *c, int nsteps) {
  if (!c || nsteps < 0) return DLBC_ERR_ARG;
  if (nsteps == 0) return DLBC_OK;
  int rc = prepare(c);
  if (rc) return rc;
  const int sched = effective_schedule(c);
  if (c->storage != DLBC_STORAGE_F64 && sched == DLBC_SCHED_FUSED) {
    // the tiled kernel reads its parameters from the device table: refresh it
    // before every replay, not only before a capture
    rc = shifted_self_ready(c);
    if (rc) return rc;
  }
  auto key = graph_key(c, sched);
  if (key != c->graph_key) {
    drop_graphs(c);
    c->graph_key.swap(key);
  }
  Timer tm{c, false};
  int t = 0;
  if (nsteps >= kGraphSteps) {
    const int par = c->popA == c->bufX ? 0 : 1;
    if (!c->graph[par]) {  // an even count: it ends on the buffers it started from
      void *const a = c->popA, *const b = c->popB;
      hipGraph_t g = nullptr;
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int rcc = 0;
      for (int k = 0; k < kGraphSteps && !rcc; ++k) rcc = step_once(c, sched, &tm);
      const hipError_t ec = hipStreamEndCapture(c->stream, &g);
      c->popA = a;  // captured, not run
      c->popB = b;
      if (rcc) {
        if (g) (void)hipGraphDestroy(g);
        return rcc;
      }
      HIPCHK(c, ec);
      const hipError_t ei = hipGraphInstantiate(&c->graph[par], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) c->graph[par] = nullptr;
      HIPCHK(c, ei);
    }
    for (; t + kGraphSteps <= nsteps; t += kGraphSteps)
      HIPCHK(c, hipGraphLaunch(c->graph[par], c->stream));
  }
  for (; t < nsteps; ++t) {
    rc = step_once(c, sched, &tm);
    if (rc) return rc;
  }
  HIPCHK(c, hipGetLastError());
  return DLBC_OK;
}

int dlbc_gpu_advect(dlbc_gpu_ctx *c) {
  if (!c) return DLBC_ERR_ARG;
  int rc = prepare(c);
  if (rc) return rc;
  rc = dispatch<AdvectOnly>(c);
  if (rc) return rc;
  HIPCHK(c, hipGetLastError());
  return DLBC_OK;
}

int dlbc_gpu_step_timed(dlbc_gpu_ctx *c, int nsteps, double *elapsed_ms, double *kernel_ms,
                        int *nkernels) {
  if (!c || nsteps < 1) return DLBC_ERR_ARG;
  int rc = prepare(c);
  if (rc) return rc;
  double acc[kMaxKernels] = {0};
  int nk = 0;
  hipEvent_t t0 = c->ev[2 * kMaxKernels], t1 = c->ev[2 * kMaxKernels + 1];
  HIPCHK(c, hipEventRecord(t0, c->stream));
  for (int t = 0; t < nsteps; ++t) {
    Timer tm{c, true};
    rc = do_steps(c, 1, &tm);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    nk = tm.slot;
    for (int k = 0; k < nk; ++k) {
      float ms = 0;
      HIPCHK(c, hipEventElapsedTime(&ms, c->ev[2 * k], c->ev[2 * k + 1]));
      acc[k] += ms;
    }
  }
  HIPCHK(c, hipEventRecord(t1, c->stream));
  HIPCHK(c, hipEventSynchronize(t1));
  float tot = 0;
  HIPCHK(c, hipEventElapsedTime(&tot, t0, t1));
  if (elapsed_ms) *elapsed_ms = tot;
  if (kernel_ms)
    for (int k = 0; k < nk; ++k) kernel_ms[k] = acc[k] / nsteps;
  if (nkernels) *nkernels = nk;
  return DLBC_OK;
}

const char *dlbc_gpu_kernel_name(const dlbc_gpu_ctx *c, int i) {
  if (!c) return "";
  const int sched = effective_schedule(c);
  if (sched == DLBC_SCHED_FUSED && c->storage != DLBC_STORAGE_F64)
    return i == 0 ? "tiled_d3q19_shifted" : "";
  if (sched == DLBC_SCHED_FUSED) return i == 0 ? "fused_d3q19" : "";
  if (sched == DLBC_SCHED_MARCH) return i == 0 ? "march_d3q19" : "";
  if (sched == DLBC_SCHED_SPLIT) {
    const bool sh = c->storage != DLBC_STORAGE_F64;
    if (c->p.sc)
      return i == 0 ? (sh ? "psi_split_shifted" : "psi_split")
                    : (i == 1 ? (sh ? "stream_force_collide_split_shifted"
                                    : "stream_force_collide_split")
                              : "");
    return i == 0 ? (sh ? "stream_collide_split_shifted" : "stream_collide_split") : "";
  }
  if (c->p.sc) return i == 0 ? "psi" : (i == 1 ? "stream_force_collide" : "");
  return i == 0 ? "stream_collide" : "";
}

int dlbc_gpu_synchronize(dlbc_gpu_ctx *c) {
  if (!c) return DLBC_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DLBC_OK;
}

int dlbc_gpu_set_schedule(dlbc_gpu_ctx *c, int schedule) {
  if (!c || schedule < 0 || schedule > 4) return DLBC_ERR_ARG;
  if (schedule == DLBC_SCHED_MARCH &&
      !(c->d == 3 && c->C == 2 && c->storage == DLBC_STORAGE_F64 &&
        [&] {
          int ty;
          bool h;
          return march_shape(c->geo, c->march_rows, c->march_hold, ty, h);
        }()))
    return fail(c, DLBC_ERR_UNSUPPORTED,
                "march schedule needs D3Q19, 2 components, FP64, n_z in {32, 64, 256, 512} "
                "and, for n_z < 512, 28 sp < 2^29");
  if (schedule == DLBC_SCHED_SPLIT && !(c->d == 3 && c->C == 2))
    return fail(c, DLBC_ERR_UNSUPPORTED, "split schedule needs D3Q19 with 2 components");
  if (schedule == DLBC_SCHED_FUSED && c->storage == DLBC_STORAGE_F64 &&
      !(c->d == 3 && c->C == 2 && fused_d3q19_supported(c->geo)))
    return fail(c, DLBC_ERR_UNSUPPORTED,
                "fused schedule needs D3Q19, 2 components, n_y %% 8 == 0 and n_z %% 32 == 0");
  if (schedule == DLBC_SCHED_TWO_PASS && c->storage != DLBC_STORAGE_F64)
    return fail(c, DLBC_ERR_UNSUPPORTED, "the two-pass schedule needs FP64 storage");
  c->schedule = schedule;
  return DLBC_OK;
}

int dlbc_gpu_get_schedule(const dlbc_gpu_ctx *c) { return c ? effective_schedule(c) : -1; }

int dlbc_gpu_set_launch(dlbc_gpu_ctx *c, int min_workgroups) {
  if (!c || min_workgroups < 0 || min_workgroups > (1 << 20)) return DLBC_ERR_ARG;
  c->target_blocks = min_workgroups;
  return DLBC_OK;
}

int dlbc_gpu_set_march(dlbc_gpu_ctx *c, int rows, int hold) {
  if (!c || rows < 0 || rows > 2 || hold < -1 || hold > 1) return DLBC_ERR_ARG;
  int ty;
  bool h;
  if (!march_shape(c->geo, rows, hold, ty, h))
    return fail(c, DLBC_ERR_UNSUPPORTED, "march schedule: rows %d, hold %d not served for n_y = "
                "%d, n_z = %d", rows, hold, c->geo.ny, c->geo.nz);
  c->march_rows = rows;
  c->march_hold = hold;
  return DLBC_OK;
}

int dlbc_gpu_set_block_order(dlbc_gpu_ctx *c, int chunk) {
  if (!c || chunk < -1 || chunk > (1 << 20)) return DLBC_ERR_ARG;
  apply_block_order(c, chunk);
  return DLBC_OK;
}

int dlbc_gpu_get_block_order(const dlbc_gpu_ctx *c) { return c ? c->geo.order_chunk : -1; }

int dlbc_gpu_density(dlbc_gpu_ctx *c, int f, double *host, int halo) {
  if (!c || !host || halo < 0 || f < 0 || f >= c->C) return DLBC_ERR_ARG;
  int rc = ensure_scratch(c, sizeof(double) * c->geo.sites);
  if (!rc) rc = field_launch(c, c->scratch, 0, f, f);
  if (!rc) rc = download_sitefield(c, c->scratch, 1, host, halo);
  return rc;
}

int dlbc_gpu_velocity(dlbc_gpu_ctx *c, int f, double *host, int halo) {
  if (!c || !host || halo < 0 || f < 0 || f >= c->C) return DLBC_ERR_ARG;
  int rc = ensure_scratch(c, sizeof(double) * c->geo.sites * c->d);
  if (!rc) rc = field_launch(c, c->scratch, 1, f, f);
  if (!rc) rc = download_sitefield(c, c->scratch, c->d, host, halo);
  return rc;
}

int dlbc_gpu_colour(dlbc_gpu_ctx *c, int f1, int f2, double *host, int halo) {
  if (!c || !host || halo < 0 || f1 < 0 || f1 >= c->C || f2 < 0 || f2 >= c->C)
    return DLBC_ERR_ARG;
  int rc = ensure_scratch(c, sizeof(double) * c->geo.sites);
  if (!rc) rc = field_launch(c, c->scratch, 2, f1, f2);
  if (!rc) rc = download_sitefield(c, c->scratch, 1, host, halo);
  return rc;
}

int dlbc_gpu_total_mass(dlbc_gpu_ctx *c, int f, double *out) {
  if (!c || !out || f < 0 || f >= c->C) return DLBC_ERR_ARG;
  return reduce_launch(c, f, 0, out);
}

int dlbc_gpu_total_mass_noalloc(dlbc_gpu_ctx *c, int f, double *out) {
  if (!c || !out || f < 0 || f >= c->C) return DLBC_ERR_ARG;
  if ((int64_t)c->geo.sites > kNoAllocMaxSites)
    return fail(c, DLBC_ERR_UNSUPPORTED,
                "total_mass_noalloc: %d sites exceed its fixed scratch (%d doubles, at most "
                "%lld sites); use dlbc_gpu_total_mass",
                c->geo.sites, kNoAllocDoubles, (long long)kNoAllocMaxSites);
  double *work = nullptr;
  {
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (c->device >= 0 && c->device < kNoAllocMaxDevices) work = noalloc_ptr[c->device];
  }
  if (!work)
    return fail(c, DLBC_ERR_STATE, "total_mass_noalloc: no scratch address for device %d",
                c->device);
  return reduce_into(c, f, 0, work, out);
}

int dlbc_gpu_total_momentum(dlbc_gpu_ctx *c, int f, double *out) {
  if (!c || !out || f < 0 || f >= c->C) return DLBC_ERR_ARG;
  for (int j = 0; j < c->d; ++j) {
    double r[2];
    int rc = reduce_launch(c, f, j + 1, r);
    if (rc) return rc;
    out[j] = r[0];
    out[c->d + j] = r[1];
  }
  return DLBC_OK;
}

int dlbc_gpu_laplace(dlbc_gpu_ctx *c, double *out) {
  if (!c || !out) return DLBC_ERR_ARG;
  if (c->C != 2 || !c->p.sc)
    return fail(c, DLBC_ERR_UNSUPPORTED, "Laplace diagnostics need two Shan-Chen components");
  if (!c->params_set || !c->pops_uploaded[0] || !c->pops_uploaded[1])
    return fail(c, DLBC_ERR_STATE, "set parameters and upload populations first");
  int cin[3], cout[3];
  for (int vd = 0; vd < c->d; ++vd) {
    cin[vd] = (int)(c->n[vd] * 0.5);  // laplace.d:73-76
    cout[vd] = 1;
  }
  double din[kMaxC], dout[kMaxC];
  int rc = cube_density(c, cin, din);
  if (!rc) rc = cube_density(c, cout, dout);
  if (rc) return rc;
  const double pin = sc_pressure(c, din), pout = sc_pressure(c, dout);
  double mass[2];
  rc = reduce_launch(c, 0, 0, mass);
  if (rc) return rc;
  const double gmass = mass[0] + mass[1];
  const double eff = gmass - (double)c->geo.sites * dout[0];  // laplace.d:95
  const double kPi = 3.14159265358979323846;
  double r, sigma;
  if (c->d == 3) {
    r = std::pow(eff / (4.0 / 3.0 * kPi * (din[0] - dout[0])), 1.0 / 3.0);
    sigma = r * (pin - pout) / 2.0;
  } else if (c->d == 2) {
    r = std::pow(eff / (kPi * (din[0] - dout[0])), 1.0 / 2.0);
    sigma = r * (pin - pout);
  } else {
    r = eff / (din[0] - dout[0]);
    sigma = r * (pin - pout);
  }
  double D = std::nan("");
  if (c->d == 2) {
    rc = ensure_scratch(c, sizeof(double) * c->geo.sites);
    if (!rc) rc = field_launch(c, c->scratch, 0, 0, 0);
    if (rc) return rc;
    std::vector<double> rho(c->geo.sites);
    std::vector<uint8_t> solid(c->geo.sites);
    HIPCHK(c, hipMemcpyAsync(rho.data(), c->scratch, sizeof(double) * c->geo.sites,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(solid.data(), c->mask, c->geo.sites, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    D = deformation_2d(c, rho, solid);
  }
  out[0] = sigma;
  out[1] = r;
  out[2] = pin;
  out[3] = pout;
  out[4] = D;
  return DLBC_OK;
}

int dlbc_gpu_eval_psi(dlbc_gpu_ctx *c, int psi_form, const double *x, double *y, int64_t n) {
  if (!c || !x || !y || n < 0 || psi_form < 0 || psi_form > 1) return DLBC_ERR_ARG;
  if (n == 0) return DLBC_OK;
  int rc = ensure_scratch(c, sizeof(double) * 2 * n);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->scratch, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_eval_psi, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, c->scratch,
                     c->scratch + n, n, psi_form);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(y, c->scratch + n, sizeof(double) * n, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DLBC_OK;
}

int64_t dlbc_selftest_site_division(int64_t n, uint64_t seed) {
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 7;
  auto next = [&]() {  // splitmix64
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  int64_t bad = 0;
  for (int64_t k = 0; k < n; ++k) {
    const uint64_t r = next();
    // divisors: every n_z or n_y n_z of a lattice (1 .. 2^30), small ones often
    uint32_t d = (k & 1) ? 1u + (uint32_t)(r % 4096u) : 1u + (uint32_t)(r % (1u << 30));
    uint64_t m;
    int l;
    geo_divisor(d, m, l);
    // dividends: site indices below 2^31, multiples of d and their neighbours often
    const uint64_t r2 = next();
    int64_t sv = (int64_t)(r2 % (1ull << 31));
    if ((k >> 1) & 1) sv = (sv / d) * d - 1 + (int64_t)(r2 >> 62);
    if (sv < 0) sv = 0;
    if (sv > 0x7fffffff) sv = 0x7fffffff;
    const int s = (int)sv;
    if (geo_div(s, m, l) != s / (int)d) ++bad;
  }
  return bad;
}

const char *dlbc_gpu_last_error(const dlbc_gpu_ctx *c) { return c ? c->err.c_str() : "null context"; }

int64_t dlbc_gpu_device_bytes(const dlbc_gpu_ctx *c) { return c ? c->dev_bytes : 0; }

void dlbc_gpu_destroy(dlbc_gpu_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graphs(c);
  if (c->popA) (void)hipFree(c->popA);
  if (c->popB) (void)hipFree(c->popB);
  if (c->S) (void)hipFree(c->S);
  if (c->P) (void)hipFree(c->P);
  if (c->mask) (void)hipFree(c->mask);
  if (c->nbw) (void)hipFree(c->nbw);
  if (c->scratch) (void)hipFree(c->scratch);
  if (c->d_self) (void)hipFree(c->d_self);
  if (c->d_items) (void)hipFree(c->d_items);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"

// ============================================================================
// Batches of lattices: one persistent launch per step over all their tiles
// ============================================================================
struct dlbc_batch {
  std::vector<dlbc_gpu_ctx *> ctx;
  int storage = 0, device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr, e0 = nullptr, e1 = nullptr;
  BatchLattice *d_lats = nullptr;
  std::vector<BatchLattice> h_lats;  // the table on the device
  bool lats_valid = false;
  BatchItem *d_items = nullptr;
  size_t items_cap = 0;
  std::vector<BatchItem> h_items;
  std::vector<int> lat_first;  // first item of each lattice, and the total
  int ty = 8, chunk = 1 << 20, grid = 512;
  int kernel = 0;  // tiled: 0 solid-capable, 1 without solids; 2: the split pair;
                   // 3: the split pair as one persistent (cooperative) launch
  int coop_grid = 0;     // workgroups of that launch (0: not yet queried)
  int coop_variant = -1;  // the kernel variant coop_grid was queried for
  int *d_first = nullptr;     // split: first 128-site block of each member, and the total
  std::vector<int> h_first;
  bool items_dirty = true;
  hipGraphExec_t graph = nullptr;
  int graph_sig = -1;  // batch_launch_sig() when the graph was captured
  std::string err;
};

namespace {

int bfail(dlbc_batch *b, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  b->err = buf;
  return code;
}

#define BHIPCHK(b, call)                                                                    \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return bfail((b), DLBC_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));       \
  } while (0)

void batch_drop_graph(dlbc_batch *b) {
  if (!b->graph) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);  // replays may be in flight
  (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
}

// Work list for the current tile height and chunk length, on the device.
int batch_build_items(dlbc_batch *b) {
  std::vector<Geo> geo;
  for (auto *c : b->ctx) geo.push_back(c->geo);
  b->h_items.clear();
  b->lat_first.assign(1, 0);
  for (size_t l = 0; l < geo.size(); ++l) {
    std::vector<BatchItem> one;
    batch_items(&geo[l], 1, b->ty, b->chunk, one);
    for (auto it : one) {
      it.lat = (int)l;
      b->h_items.push_back(it);
    }
    b->lat_first.push_back((int)b->h_items.size());
  }
  if (b->h_items.size() > b->items_cap) {
    if (b->d_items) BHIPCHK(b, hipFree(b->d_items));
    b->d_items = nullptr;
    BHIPCHK(b, hipMalloc(&b->d_items, sizeof(BatchItem) * b->h_items.size()));
    b->items_cap = b->h_items.size();
  }
  BHIPCHK(b, hipMemcpyAsync(b->d_items, b->h_items.data(), sizeof(BatchItem) * b->h_items.size(),
                            hipMemcpyHostToDevice, b->stream));
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  b->items_dirty = false;
  batch_drop_graph(b);
  return DLBC_OK;
}

bool batch_has_solid(const dlbc_batch *b) {
  for (auto *c : b->ctx)
    if (c->has_solid) return true;
  return false;
}

bool batch_has_sc(const dlbc_batch *b, int l0, int l1) {
  for (int l = l0; l < l1; ++l)
    if (b->ctx[l]->sparams.sc) return true;
  return false;
}

// The host state that picks kernels at launch beyond the configuration: any
// member with solids (the tiled kernel's SOLID variant) and any with
// Shan-Chen (the split schedule's psi pass).  A captured graph is only valid
// for the value it was captured with.
int batch_launch_sig(const dlbc_batch *b) {
  return (batch_has_solid(b) ? 1 : 0) | (batch_has_sc(b, 0, (int)b->ctx.size()) ? 2 : 0);
}

// Make every member ready (solid densities, pending uploads done) and the
// device table current (buffers swap between calls, parameters may change).
int batch_prepare(dlbc_batch *b) {
  std::vector<BatchLattice> t(b->ctx.size());
  for (size_t l = 0; l < b->ctx.size(); ++l) {
    dlbc_gpu_ctx *c = b->ctx[l];
    int rc = prepare(c);
    if (rc) return bfail(b, rc, "lattice %zu: %s", l, c->err.c_str());
    BHIPCHK(b, hipStreamSynchronize(c->stream));
    std::memset(&t[l], 0, sizeof(BatchLattice));
    t[l].pop[0] = c->popA;
    t[l].pop[1] = c->popB;
    t[l].nbw = c->nbw;
    t[l].S = c->S;
    t[l].P = c->P;
    t[l].g = c->geo;
    t[l].p = c->sparams;
  }
  if (!b->lats_valid || std::memcmp(t.data(), b->h_lats.data(), sizeof(BatchLattice) * t.size())) {
    BHIPCHK(b, hipMemcpyAsync(b->d_lats, t.data(), sizeof(BatchLattice) * t.size(),
                              hipMemcpyHostToDevice, b->stream));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    b->h_lats = t;
    b->lats_valid = true;
  }
  if (b->graph && b->graph_sig != batch_launch_sig(b)) batch_drop_graph(b);
  if (b->items_dirty) return batch_build_items(b);
  return DLBC_OK;
}

// Whether every member's element offsets fit the split kernels' 32-bit lane
// offsets (the OFF32 variants).
bool batch_off32(const dlbc_batch *b) {
  for (auto *c : b->ctx)
    if (!split_off32(c->geo)) return false;
  return true;
}

// Split schedule over the members [l0, l1).
int batch_launch_split(dlbc_batch *b, int l0, int l1, int parity) {
  const int nlat = (int)b->ctx.size();
  const int blk0 = b->h_first[l0], nblk = b->h_first[l1] - blk0;
  const bool psi = batch_has_sc(b, l0, l1);
  const bool f32 = b->storage == DLBC_STORAGE_F32_SHIFT, o32 = batch_off32(b);
  if (f32 && o32)
    launch_batch_split<float, true>(b->d_lats, b->d_first, nlat, blk0, nblk, parity, psi,
                                    b->stream);
  else if (f32)
    launch_batch_split<float, false>(b->d_lats, b->d_first, nlat, blk0, nblk, parity, psi,
                                     b->stream);
  else if (o32)
    launch_batch_split<__half, true>(b->d_lats, b->d_first, nlat, blk0, nblk, parity, psi,
                                     b->stream);
  else
    launch_batch_split<__half, false>(b->d_lats, b->d_first, nlat, blk0, nblk, parity, psi,
                                      b->stream);
  BHIPCHK(b, hipGetLastError());
  return DLBC_OK;
}

// The split pair over every member as one persistent launch.
int batch_launch_persistent(dlbc_batch *b, int parity) {
  const int nlat = (int)b->ctx.size(), nblk = b->h_first[nlat];
  if (nblk <= 0) return DLBC_OK;
  const bool psi = batch_has_sc(b, 0, nlat);
  const bool f32 = b->storage == DLBC_STORAGE_F32_SHIFT, o32 = batch_off32(b);
  // the kernel (and so its occupancy) follows the storage and the offset width
  const int variant = (f32 ? 2 : 0) | (o32 ? 1 : 0);
  if (b->coop_grid <= 0 || b->coop_variant != variant) {
    b->coop_grid = f32 ? (o32 ? batch_split_persistent_grid<float, true>(b->device, nblk)
                              : batch_split_persistent_grid<float, false>(b->device, nblk))
                       : (o32 ? batch_split_persistent_grid<__half, true>(b->device, nblk)
                              : batch_split_persistent_grid<__half, false>(b->device, nblk));
    if (b->coop_grid <= 0)
      return bfail(b, DLBC_ERR_HIP, "persistent split: occupancy query failed");
    b->coop_variant = variant;
  }
  const int grid = std::min(b->coop_grid, nblk);
  hipError_t e;
  if (f32 && o32)
    e = launch_batch_split_persistent<float, true>(b->d_lats, b->d_first, nlat, nblk, parity,
                                                   psi, grid, b->stream);
  else if (f32)
    e = launch_batch_split_persistent<float, false>(b->d_lats, b->d_first, nlat, nblk, parity,
                                                    psi, grid, b->stream);
  else if (o32)
    e = launch_batch_split_persistent<__half, true>(b->d_lats, b->d_first, nlat, nblk, parity,
                                                    psi, grid, b->stream);
  else
    e = launch_batch_split_persistent<__half, false>(b->d_lats, b->d_first, nlat, nblk, parity,
                                                     psi, grid, b->stream);
  BHIPCHK(b, e);
  return DLBC_OK;
}

int batch_launch(dlbc_batch *b, const BatchItem *items, int nitems, int parity, int grid) {
  if (b->kernel == 3) return batch_launch_persistent(b, parity);
  if (b->kernel == 2) return batch_launch_split(b, 0, (int)b->ctx.size(), parity);
  const bool solid_kernel = b->kernel == 0 || batch_has_solid(b);
  const int rc = b->storage == DLBC_STORAGE_F32_SHIFT
                     ? launch_batch_d3q19<float>(b->d_lats, items, nitems, parity, b->ty, grid,
                                                 solid_kernel, b->stream)
                     : launch_batch_d3q19<__half>(b->d_lats, items, nitems, parity, b->ty, grid,
                                                  solid_kernel, b->stream);
  if (rc) return bfail(b, DLBC_ERR_ARG, "unsupported tile height %d", b->ty);
  BHIPCHK(b, hipGetLastError());
  return DLBC_OK;
}

// After nsteps: members whose buffers advanced an odd number of steps swap
// them, and later work on a member's own stream waits for the batch.
int batch_finish(dlbc_batch *b, int nsteps) {
  BHIPCHK(b, hipEventRecord(b->done, b->stream));
  for (auto *c : b->ctx) {
    if (nsteps & 1) std::swap(c->popA, c->popB);
    BHIPCHK(b, hipStreamWaitEvent(c->stream, b->done, 0));
  }
  return DLBC_OK;
}

}  // namespace

extern "C" {

int dlbc_batch_create(dlbc_batch **out, dlbc_gpu_ctx *const *lattices, int n) {
  if (!out || !lattices || n < 1) return DLBC_ERR_ARG;
  *out = nullptr;
  const int storage = lattices[0] ? lattices[0]->storage : -1;
  for (int l = 0; l < n; ++l) {
    const dlbc_gpu_ctx *c = lattices[l];
    if (!c) return DLBC_ERR_ARG;
    if (c->d != 3 || c->q != 19 || c->C != 2 || c->storage == DLBC_STORAGE_F64 ||
        c->storage != storage || c->device != lattices[0]->device)
      return DLBC_ERR_UNSUPPORTED;
    for (int k = 0; k < l; ++k)
      if (lattices[k] == c) return DLBC_ERR_ARG;  // a lattice may appear once
  }
  auto *b = new dlbc_batch();
  b->ctx.assign(lattices, lattices + n);
  b->storage = storage;
  b->device = lattices[0]->device;
  auto bail = [&](int code) {
    dlbc_batch_destroy(b);
    return code;
  };
  if (hipSetDevice(b->device) != hipSuccess) return bail(DLBC_ERR_HIP);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(DLBC_ERR_HIP);
  if (hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess ||
      hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess)
    return bail(DLBC_ERR_HIP);
  if (hipMalloc(&b->d_lats, sizeof(BatchLattice) * n) != hipSuccess) return bail(DLBC_ERR_NOMEM);
  b->h_first.assign(1, 0);
  for (int l = 0; l < n; ++l)
    b->h_first.push_back(b->h_first.back() + shift_grid(lattices[l]->geo.sites));
  if (hipMalloc(&b->d_first, sizeof(int) * (n + 1)) != hipSuccess) return bail(DLBC_ERR_NOMEM);
  if (hipMemcpy(b->d_first, b->h_first.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice) !=
      hipSuccess)
    return bail(DLBC_ERR_HIP);
  *out = b;
  return DLBC_OK;
}

int dlbc_batch_set_config(dlbc_batch *b, int tile_y, int chunk, int workgroups, int kernel) {
  if (!b || (tile_y != 4 && tile_y != 8) || chunk < 4 || workgroups < 8 || kernel < 0 ||
      kernel > 3)
    return DLBC_ERR_ARG;
  if (tile_y != b->ty || chunk != b->chunk) b->items_dirty = true;
  if (workgroups != b->grid || kernel != b->kernel) batch_drop_graph(b);
  b->ty = tile_y;
  b->chunk = chunk;
  b->grid = workgroups;
  b->kernel = kernel;
  return DLBC_OK;
}

int dlbc_batch_get_config(const dlbc_batch *b, int *out) {
  if (!b || !out) return DLBC_ERR_ARG;
  out[0] = b->ty;
  out[1] = b->chunk;
  out[2] = b->grid;
  out[3] = b->kernel;
  return DLBC_OK;
}

int dlbc_batch_step(dlbc_batch *b, int nsteps, int use_graph) {
  if (!b || nsteps < 0) return DLBC_ERR_ARG;
  if (nsteps == 0) return DLBC_OK;
  int rc = batch_prepare(b);
  if (rc) return rc;
  const int nit = (int)b->h_items.size();
  int t = 0;
  // the persistent split launch is cooperative and is not captured: one
  // launch per step either way
  if (use_graph && nsteps >= 2 && b->kernel != 3) {
    if (!b->graph) {  // two steps: even parity, then odd
      hipGraph_t g = nullptr;
      BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc0 = batch_launch(b, b->d_items, nit, 0, b->grid);
      int rc1 = rc0 ? rc0 : batch_launch(b, b->d_items, nit, 1, b->grid);
      const hipError_t ec = hipStreamEndCapture(b->stream, &g);
      if (rc1) {
        if (g) (void)hipGraphDestroy(g);
        return rc1;
      }
      BHIPCHK(b, ec);
      const hipError_t ei = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      BHIPCHK(b, ei);
      b->graph_sig = batch_launch_sig(b);
    }
    for (; t + 2 <= nsteps; t += 2) BHIPCHK(b, hipGraphLaunch(b->graph, b->stream));
  }
  for (; t < nsteps; ++t) {
    rc = batch_launch(b, b->d_items, nit, t & 1, b->grid);
    if (rc) return rc;
  }
  return batch_finish(b, nsteps);
}

int dlbc_batch_step_looped(dlbc_batch *b, int nsteps) {
  if (!b || nsteps < 0) return DLBC_ERR_ARG;
  if (nsteps == 0) return DLBC_OK;
  int rc = batch_prepare(b);
  if (rc) return rc;
  for (int t = 0; t < nsteps; ++t)
    for (size_t l = 0; l + 1 < b->lat_first.size(); ++l) {
      const int first = b->lat_first[l], cnt = b->lat_first[l + 1] - first;
      rc = b->kernel >= 2 ? batch_launch_split(b, (int)l, (int)l + 1, t & 1)
                          : batch_launch(b, b->d_items + first, cnt, t & 1,
                                         std::min(b->grid, (cnt + 7) / 8 * 8));
      if (rc) return rc;
    }
  return batch_finish(b, nsteps);
}

int dlbc_batch_autotune(dlbc_batch *b, int *out, double *best_ms) {
  if (!b) return DLBC_ERR_ARG;
  int rc = batch_prepare(b);
  if (rc) return rc;
  const int tys[2] = {8, 4};
  const int chunks[4] = {16, 32, 64, 1 << 20};
  const int grids[4] = {256, 512, 1024, 2048};
  double best = 1e300;
  int bt = b->ty, bc = b->chunk, bg = b->grid, bk = b->kernel;
  const int nkernels = batch_has_solid(b) ? 1 : 2;
  auto trial = [&](int nit, int gr, double &ms_out) -> int {
    // parity 0 only: reads the current state, writes the other buffer (and psi)
    int r = batch_launch(b, b->d_items, nit, 0, gr);
    if (r) return r;
    BHIPCHK(b, hipEventRecord(b->e0, b->stream));
    for (int k = 0; k < 3; ++k) {
      r = batch_launch(b, b->d_items, nit, 0, gr);
      if (r) return r;
    }
    BHIPCHK(b, hipEventRecord(b->e1, b->stream));
    BHIPCHK(b, hipEventSynchronize(b->e1));
    float ms = 0.0f;
    BHIPCHK(b, hipEventElapsedTime(&ms, b->e0, b->e1));
    ms_out = ms / 3.0;
    return DLBC_OK;
  };
  for (int kn : {2, 3}) {  // the split pair, two launches or one: no tile or chunk to tune
    b->kernel = kn;
    double ms;
    rc = trial(0, b->grid, ms);
    if (rc) return rc;
    if (ms < best) {
      best = ms;
      bk = kn;
    }
  }
  for (int ty : tys)
    for (int ch : chunks) {
      b->ty = ty;
      b->chunk = ch;
      rc = batch_build_items(b);
      if (rc) return rc;
      const int nit = (int)b->h_items.size();
      for (int kn = 0; kn < nkernels; ++kn)
      for (int gr : grids) {
        b->kernel = kn;
        double ms;
        rc = trial(nit, gr, ms);
        if (rc) return rc;
        if (ms < best) {
          best = ms;
          bt = ty;
          bc = ch;
          bg = gr;
          bk = kn;
        }
      }
    }
  b->ty = bt;
  b->chunk = bc;
  b->grid = bg;
  b->kernel = bk;
  rc = batch_build_items(b);
  if (rc) return rc;
  if (out) {
    out[0] = bt;
    out[1] = bc;
    out[2] = bg;
    out[3] = bk;
  }
  if (best_ms) *best_ms = best;
  return DLBC_OK;
}

int dlbc_batch_synchronize(dlbc_batch *b) {
  if (!b) return DLBC_ERR_ARG;
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  return DLBC_OK;
}

const char *dlbc_batch_last_error(const dlbc_batch *b) { return b ? b->err.c_str() : ""; }

void dlbc_batch_destroy(dlbc_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  if (b->d_lats) (void)hipFree(b->d_lats);
  if (b->d_items) (void)hipFree(b->d_items);
  if (b->d_first) (void)hipFree(b->d_first);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

}  // extern "C"
