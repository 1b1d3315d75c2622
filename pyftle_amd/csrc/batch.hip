// Batched FTLE windows: one persistent launch over a work list of
// (field, block of seeds) items drawn from a device counter.  The per-item
// work is the cell-local FP32 window of advect_impl.hpp with its FTLE
// epilogue; each field brings its own grid, snapshot ring, seeds, step count,
// step size and time origin through a descriptor table in device memory.
#include <stddef.h>

#include <algorithm>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "advect_impl.hpp"

struct pyftle_batch {
  int dim = 0, method = 0, tmode = 0, pair = 0, cell = 0;
  int threads = 0, grid = 0;
  int64_t n_items = 0;
  pyftle::AdvectArgs *d_table = nullptr;
  int2 *d_items = nullptr;
  unsigned int *d_counter = nullptr;
  bool use_graph = false;
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace pyftle {

__global__ void reset_counter_kernel(unsigned int *c) { *c = 0u; }

template <int DIM, int METHOD, int TMODE, bool PAIR, bool CELL>
__global__ void __launch_bounds__(512) batch_kernel(const AdvectArgs *__restrict__ table,
                                                    const int2 *__restrict__ items, int n_items,
                                                    unsigned int *counter) {
  __shared__ int s_item;
  for (;;) {
    __syncthreads();  // the previous item (and its LDS epilogue) is finished
    if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= n_items) break;
    const int f = __builtin_amdgcn_readfirstlane(items[it].x);
    const int b = __builtin_amdgcn_readfirstlane(items[it].y);
    advect_block_cell<DIM, METHOD, TMODE, PAIR, true, CELL>(table[f], b);
  }
}

// The same persistent loop with the work list implied by a prefix sum of the
// fields' block counts (fields stored in processing order): item `it` is block
// it - prefix[f] of the field f with prefix[f] <= it < prefix[f + 1].  Used by
// the allocation-free entry point, whose scratch holds no item table.
template <int DIM, int METHOD, int TMODE, bool PAIR, bool CELL>
__global__ void __launch_bounds__(512) batch_kernel_prefix(const AdvectArgs *__restrict__ table,
                                                           const int64_t *__restrict__ prefix,
                                                           int n_fields, int n_items,
                                                           unsigned int *counter) {
  __shared__ int s_item;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= n_items) break;
    int lo = 0, hi = n_fields - 1;  // largest f with prefix[f] <= it
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= it) lo = mid;
      else hi = mid - 1;
    }
    const int f = __builtin_amdgcn_readfirstlane(lo);
    const int b = __builtin_amdgcn_readfirstlane((int)(it - prefix[f]));
    advect_block_cell<DIM, METHOD, TMODE, PAIR, true, CELL>(table[f], b);
  }
}

// Counting copy of batch_kernel (RK4 windows on per-cell records): the same items, the
// same per-item device function and results, and every wave adds the 16-byte record loads
// it issues to *loads.  Run outside timed regions.
template <int DIM, int TMODE, bool PAIR>
__global__ void __launch_bounds__(512) batch_count_kernel(const AdvectArgs *__restrict__ table,
                                                          const int2 *__restrict__ items, int n_items,
                                                          unsigned int *counter, unsigned long long *loads) {
  __shared__ int s_item;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= n_items) break;
    const int f = __builtin_amdgcn_readfirstlane(items[it].x);
    const int b = __builtin_amdgcn_readfirstlane(items[it].y);
    AdvectArgs a = table[f];
    a.load_count = loads;
    advect_block_cell<DIM, PYFTLE_RK4, TMODE, PAIR, true, true, true>(a, b);
  }
}

template <int DIM>
static const void *count_kernel_for(int tmode, int pair) {
  if (tmode == PYFTLE_TIME_LINEAR) return reinterpret_cast<const void *>(&batch_count_kernel<DIM, PYFTLE_TIME_LINEAR, true>);
  if (tmode == PYFTLE_TIME_FROZEN)
    return pair ? reinterpret_cast<const void *>(&batch_count_kernel<DIM, PYFTLE_TIME_FROZEN, true>)
                : reinterpret_cast<const void *>(&batch_count_kernel<DIM, PYFTLE_TIME_FROZEN, false>);
  return pair ? reinterpret_cast<const void *>(&batch_count_kernel<DIM, PYFTLE_TIME_STEADY, true>)
              : reinterpret_cast<const void *>(&batch_count_kernel<DIM, PYFTLE_TIME_STEADY, false>);
}

using BatchFn = void (*)(const AdvectArgs *, const int2 *, int, unsigned int *);

template <int DIM, int METHOD, int TMODE, bool PAIR>
static const void *kernel_ptr(int cell, bool prefix) {
  if (prefix)
    return cell ? reinterpret_cast<const void *>(&batch_kernel_prefix<DIM, METHOD, TMODE, PAIR, true>)
                : reinterpret_cast<const void *>(&batch_kernel_prefix<DIM, METHOD, TMODE, PAIR, false>);
  return cell ? reinterpret_cast<const void *>(&batch_kernel<DIM, METHOD, TMODE, PAIR, true>)
              : reinterpret_cast<const void *>(&batch_kernel<DIM, METHOD, TMODE, PAIR, false>);
}

template <int DIM, int METHOD>
static const void *pick_time(int tmode, int pair, int cell, bool prefix) {
  if (tmode == PYFTLE_TIME_LINEAR) return kernel_ptr<DIM, METHOD, PYFTLE_TIME_LINEAR, true>(cell, prefix);
  if (tmode == PYFTLE_TIME_FROZEN)
    return pair ? kernel_ptr<DIM, METHOD, PYFTLE_TIME_FROZEN, true>(cell, prefix)
                : kernel_ptr<DIM, METHOD, PYFTLE_TIME_FROZEN, false>(cell, prefix);
  return pair ? kernel_ptr<DIM, METHOD, PYFTLE_TIME_STEADY, true>(cell, prefix)
              : kernel_ptr<DIM, METHOD, PYFTLE_TIME_STEADY, false>(cell, prefix);
}

template <int DIM>
static const void *pick_method(int method, int tmode, int pair, int cell, bool prefix) {
  if (method == PYFTLE_RK4) return pick_time<DIM, PYFTLE_RK4>(tmode, pair, cell, prefix);
  if (method == PYFTLE_EULER) return pick_time<DIM, PYFTLE_EULER>(tmode, pair, cell, prefix);
  return pick_time<DIM, PYFTLE_AB2>(tmode, pair, cell, prefix);
}

static const void *batch_kernel_for(int dim, int method, int tmode, int pair, int cell, bool prefix) {
  return dim == 2 ? pick_method<2>(method, tmode, pair, cell, prefix)
                  : pick_method<3>(method, tmode, pair, cell, prefix);
}

static const void *batch_kernel_for(const pyftle_batch *b) {
  return batch_kernel_for(b->dim, b->method, b->tmode, b->pair, b->cell, false);
}

// Checks the options and every field of a batch and fills one descriptor per
// field (table[i]) and its block count (blocks[i]); shared by both entry points.
struct BatchShape {
  int dim = 0, layout = 0, nb = 0, G = 0, threads = 0;
  int64_t n_items = 0;
};

static int prepare_batch(const char *who, const pyftle_batch_field *fields, int32_t n_fields,
                         const pyftle_batch_opts *opts, AdvectArgs *table, int64_t *blocks,
                         BatchShape *shape) {
  if (!fields || n_fields <= 0) return fail(std::string(who) + ": no fields");
  if (!opts) return fail(std::string(who) + ": opts is NULL");
  if (opts->method != PYFTLE_EULER && opts->method != PYFTLE_AB2 && opts->method != PYFTLE_RK4)
    return fail(std::string(who) + ": unknown integrator");
  if (opts->time_mode != PYFTLE_TIME_STEADY && opts->time_mode != PYFTLE_TIME_FROZEN &&
      opts->time_mode != PYFTLE_TIME_LINEAR)
    return fail(std::string(who) + ": unknown time mode");
  if (opts->pos_dtype != PYFTLE_F32 && opts->pos_dtype != PYFTLE_F64)
    return fail(std::string(who) + ": bad positions dtype");
  if (opts->lane_order != PYFTLE_ORDER_SEED_MAJOR && opts->lane_order != PYFTLE_ORDER_NEIGHBOR_MAJOR)
    return fail(std::string(who) + ": unknown lane order");
  const int dim = fields[0].field.dim;
  const int layout = fields[0].field.layout;
  const int nb = 2 * dim;
  const int G = opts->groups_per_block > 0 ? opts->groups_per_block : 64;
  if (G * nb > 512) return fail(std::string(who) + ": groups_per_block * 2*dim must be <= 512");
  int64_t n_items = 0;
  for (int i = 0; i < n_fields; ++i) {
    const pyftle_batch_field &bf = fields[i];
    int rc = make_field_dev(&bf.field, &table[i].f);
    if (rc) return fail(std::string(who) + ": field " + std::to_string(i) + ": " + pyftle_last_error(), rc);
    if (bf.field.dim != dim || bf.field.layout != layout)
      return fail(std::string(who) + ": all fields must share dim and record layout");
    if (bf.field.dtype != PYFTLE_F32) return fail(std::string(who) + ": batches run FP32 fields");
    if (bf.field.interp != PYFTLE_INTERP_GINTERP)
      return fail(std::string(who) + ": batches run GridInterp-mode fields (interp 0)");
    if (opts->time_mode == PYFTLE_TIME_LINEAR &&
        (!layout_pair(layout) || !(bf.field.dt_snap != 0.0)))
      return fail(std::string(who) + ": LINEAR time needs PAIR or CELL_PAIR fields with non-zero dt_snap");
    if (opts->time_mode == PYFTLE_TIME_FROZEN &&
        (bf.first_snapshot < 0 || (int64_t)bf.first_snapshot + bf.n_steps > bf.field.n_snapshots))
      return fail(std::string(who) + ": FROZEN steps run past the last snapshot of field " + std::to_string(i));
    if (opts->time_mode == PYFTLE_TIME_STEADY &&
        (bf.first_snapshot < 0 || bf.first_snapshot >= bf.field.n_snapshots))
      return fail(std::string(who) + ": STEADY snapshot out of range in field " + std::to_string(i));
    if (bf.n_groups < 0 || bf.n_steps < 0) return fail(std::string(who) + ": negative size");
    if (bf.n_groups > 0 && (!bf.positions_in || !bf.ftle_out))
      return fail(std::string(who) + ": NULL positions or output in field " + std::to_string(i));
    AdvectArgs &a = table[i];
    a.pos_in = bf.positions_in;
    a.pos_out = bf.positions_out;
    a.vprev = nullptr;
    a.ftle_out = bf.ftle_out;
    a.n_groups = bf.n_groups;
    a.nb = nb;
    a.G = G;
    a.lane_order = opts->lane_order;
    a.remap = 0;
    a.pos_f64 = opts->pos_dtype == PYFTLE_F64;
    a.n_steps = bf.n_steps;
    a.first = bf.first_snapshot;
    a.have_prev = 0;
    a.t_start = bf.t_start;
    a.h = bf.h;
    a.map_period = bf.map_period;
    a.order = bf.seed_order;
    blocks[i] = (bf.n_groups + G - 1) / G;
    n_items += blocks[i];
  }
  if (n_items >= INT32_MAX) return fail(std::string(who) + ": too many work items");
  shape->dim = dim;
  shape->layout = layout;
  shape->nb = nb;
  shape->G = G;
  shape->threads = ((G * nb + 63) / 64) * 64;
  shape->n_items = n_items;
  return PYFTLE_OK;
}

// persistent work-groups: every CU at its occupancy limit (or blocks_per_cu), at most one per item
static int persistent_grid(const void *kernel, int threads, int blocks_per_cu, int64_t n_items,
                           int *grid, std::string *err) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0);
  if (e != hipSuccess) {
    *err = hipGetErrorString(e);
    return PYFTLE_EHIP;
  }
  per_cu = std::max(1, per_cu);
  if (blocks_per_cu > 0) per_cu = std::min(per_cu, blocks_per_cu);
  *grid = std::max(1, (int)std::min<int64_t>((int64_t)cus * per_cu, std::max<int64_t>(n_items, 1)));
  return PYFTLE_OK;
}

static int enqueue(pyftle_batch *b, hipStream_t st) {
  reset_counter_kernel<<<1, 1, 0, st>>>(b->d_counter);
  const AdvectArgs *table = b->d_table;
  const int2 *items = b->d_items;
  int n = (int)b->n_items;
  unsigned int *counter = b->d_counter;
  void *args[] = {(void *)&table, (void *)&items, (void *)&n, (void *)&counter};
  hipError_t e = hipLaunchKernel(batch_kernel_for(b), dim3(b->grid), dim3(b->threads), args, 0, st);
  if (e != hipSuccess) return fail(std::string("batch launch: ") + hipGetErrorString(e), PYFTLE_EHIP);
  return check_launch("batch launch");
}

static void release(pyftle_batch *b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  if (b->d_table) (void)hipFree(b->d_table);
  if (b->d_items) (void)hipFree(b->d_items);
  if (b->d_counter) (void)hipFree(b->d_counter);
  delete b;
}

// ---- allocation-free batches ----
// The descriptor table, the prefix sums of block counts and the work counter of
// pyftle_batch_run_noalloc live in one static device array per GPU (`scratch`),
// laid out as NoallocScratch.  Host code never passes the array's name to a
// kernel or copy: its device address is looked up once per device with
// hipGetSymbolAddress and kept in g_scratch_addr.
struct NoallocScratch {
  unsigned int counter;
  unsigned int pad_[63];
  int64_t prefix[PYFTLE_BATCH_NOALLOC_MAX_FIELDS + 1];
  alignas(16) AdvectArgs table[PYFTLE_BATCH_NOALLOC_MAX_FIELDS];
};
static_assert(sizeof(NoallocScratch) <= PYFTLE_BATCH_NOALLOC_SCRATCH_BYTES,
              "PYFTLE_BATCH_NOALLOC_SCRATCH_BYTES is smaller than the scratch layout");
static_assert(offsetof(NoallocScratch, prefix) % 16 == 0 && offsetof(NoallocScratch, table) % 16 == 0,
              "scratch sections must stay 16-byte aligned");

static __device__ __align__(256) unsigned char scratch[PYFTLE_BATCH_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mutex;                 // guards the host arrays below
static NoallocScratch *g_scratch_addr[kMaxDevices];  // device address of `scratch` per device
static NoallocScratch g_stage;                     // host staging of prefix + table (no heap)
static AdvectArgs g_table[PYFTLE_BATCH_NOALLOC_MAX_FIELDS];  // descriptors in field order
static int64_t g_blocks[PYFTLE_BATCH_NOALLOC_MAX_FIELDS];
static int g_order[PYFTLE_BATCH_NOALLOC_MAX_FIELDS];

// device address of this device's `scratch`, looked up on the first call on each device
static int scratch_address(int dev, NoallocScratch **out) {
  if (dev < 0 || dev >= kMaxDevices) return fail("batch_run_noalloc: device index beyond 64");
  if (!g_scratch_addr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(std::string("batch_run_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e),
                  PYFTLE_EHIP);
    g_scratch_addr[dev] = static_cast<NoallocScratch *>(p);
  }
  *out = g_scratch_addr[dev];
  return PYFTLE_OK;
}

}  // namespace pyftle

using namespace pyftle;

extern "C" {

int pyftle_batch_create(const pyftle_batch_field *fields, int32_t n_fields,
                        const pyftle_batch_opts *opts, pyftle_batch **out) {
  if (!out) return fail("batch_create: out is NULL");
  *out = nullptr;
  std::vector<AdvectArgs> table(std::max(n_fields, 1));
  std::vector<int64_t> blocks(std::max(n_fields, 1));
  BatchShape shape;
  int rc = prepare_batch("batch_create", fields, n_fields, opts, table.data(), blocks.data(), &shape);
  if (rc) return rc;
  const int64_t n_items = shape.n_items;
  std::vector<int> order(n_fields);
  std::iota(order.begin(), order.end(), 0);
  if (opts->longest_first)
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return fields[x].n_steps > fields[y].n_steps; });
  std::vector<int2> items;
  items.reserve((size_t)n_items);
  for (int i : order)
    for (int64_t k = 0; k < blocks[i]; ++k) items.push_back(make_int2(i, (int)k));

  pyftle_batch *b = new pyftle_batch();
  b->dim = shape.dim;
  b->method = opts->method;
  b->tmode = opts->time_mode;
  b->pair = layout_pair(shape.layout);
  b->cell = layout_cell(shape.layout);
  b->threads = shape.threads;
  b->n_items = n_items;
  b->use_graph = opts->use_graph != 0;
  auto hip_fail = [&](const char *what, hipError_t e) {
    release(b);
    return fail(std::string("batch_create: ") + what + ": " + hipGetErrorString(e), PYFTLE_EHIP);
  };
  std::string gerr;
  if (persistent_grid(batch_kernel_for(b), b->threads, opts->blocks_per_cu, n_items, &b->grid, &gerr)) {
    release(b);
    return fail("batch_create: launch shape: " + gerr, PYFTLE_EHIP);
  }
  hipError_t e = hipMalloc(&b->d_table, sizeof(AdvectArgs) * n_fields);
  if (e != hipSuccess) return hip_fail("hipMalloc table", e);
  e = hipMalloc(&b->d_items, sizeof(int2) * std::max<int64_t>(n_items, 1));
  if (e != hipSuccess) return hip_fail("hipMalloc items", e);
  e = hipMalloc(&b->d_counter, sizeof(unsigned int));
  if (e != hipSuccess) return hip_fail("hipMalloc counter", e);
  e = hipMemcpy(b->d_table, table.data(), sizeof(AdvectArgs) * n_fields, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail("copy table", e);
  if (n_items > 0) {
    e = hipMemcpy(b->d_items, items.data(), sizeof(int2) * n_items, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail("copy items", e);
  }
  if (b->use_graph && n_items > 0) {
    e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail("stream", e);
    e = hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail("begin capture", e);
    int rc = enqueue(b, b->cap_stream);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->cap_stream, &g);
    b->graph = g;
    if (rc) {
      release(b);
      return rc;
    }
    if (e != hipSuccess) return hip_fail("end capture", e);
    e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return hip_fail("instantiate", e);
  }
  *out = b;
  return PYFTLE_OK;
}

int pyftle_batch_launch(pyftle_batch *b, void *stream) {
  if (!b) return fail("batch_launch: NULL batch");
  if (b->n_items == 0) return PYFTLE_OK;
  hipStream_t st = as_stream(stream);
  if (b->use_graph) {
    hipError_t e = hipGraphLaunch(b->exec, st);
    if (e != hipSuccess) return fail(std::string("batch graph launch: ") + hipGetErrorString(e), PYFTLE_EHIP);
    return PYFTLE_OK;
  }
  return enqueue(b, st);
}

int pyftle_batch_count_loads(pyftle_batch *b, unsigned long long *loads, void *stream) {
  if (!b) return fail("batch_count_loads: NULL batch");
  if (!loads) return fail("batch_count_loads: loads is NULL");
  if (!b->cell || b->method != PYFTLE_RK4)
    return fail("batch_count_loads: counts RK4 batches on CELL / CELL_PAIR records");
  if (b->n_items == 0) return PYFTLE_OK;
  hipStream_t st = as_stream(stream);
  reset_counter_kernel<<<1, 1, 0, st>>>(b->d_counter);
  const AdvectArgs *table = b->d_table;
  const int2 *items = b->d_items;
  int n = (int)b->n_items;
  unsigned int *counter = b->d_counter;
  void *args[] = {(void *)&table, (void *)&items, (void *)&n, (void *)&counter, (void *)&loads};
  const void *k = b->dim == 2 ? count_kernel_for<2>(b->tmode, b->pair) : count_kernel_for<3>(b->tmode, b->pair);
  hipError_t e = hipLaunchKernel(k, dim3(b->grid), dim3(b->threads), args, 0, st);
  if (e != hipSuccess) return fail(std::string("batch_count_loads: ") + hipGetErrorString(e), PYFTLE_EHIP);
  return check_launch("batch_count_loads");
}

int pyftle_batch_info(const pyftle_batch *b, int64_t *n_items, int32_t *grid_blocks,
                      int32_t *block_threads) {
  if (!b) return fail("batch_info: NULL batch");
  if (n_items) *n_items = b->n_items;
  if (grid_blocks) *grid_blocks = b->grid;
  if (block_threads) *block_threads = b->threads;
  return PYFTLE_OK;
}

int pyftle_batch_destroy(pyftle_batch *b) {
  if (!b) return PYFTLE_OK;
  // a replay of this plan may still be running: its graph, table and counter are freed below
  hipError_t e = hipDeviceSynchronize();
  release(b);
  if (e != hipSuccess) return fail(std::string("batch_destroy: ") + hipGetErrorString(e), PYFTLE_EHIP);
  return PYFTLE_OK;
}

int pyftle_batch_run_noalloc(const pyftle_batch_field *fields, int32_t n_fields,
                             const pyftle_batch_opts *opts, void *stream) {
  if (n_fields > PYFTLE_BATCH_NOALLOC_MAX_FIELDS)
    return fail("batch_run_noalloc: at most " + std::to_string(PYFTLE_BATCH_NOALLOC_MAX_FIELDS) +
                " fields per call (got " + std::to_string(n_fields) + ")");
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  BatchShape shape;
  int rc = prepare_batch("batch_run_noalloc", fields, n_fields, opts, g_table, g_blocks, &shape);
  if (rc) return rc;
  if (shape.n_items == 0) return PYFTLE_OK;
  // processing order: stable, longest first (as pyftle_batch_create); the staged
  // table holds the fields in that order and prefix[] counts the blocks before each
  for (int i = 0; i < n_fields; ++i) g_order[i] = i;
  if (opts->longest_first)  // stable insertion sort (no temporary buffer)
    for (int i = 1; i < n_fields; ++i) {
      const int v = g_order[i];
      int j = i;
      for (; j > 0 && fields[g_order[j - 1]].n_steps < fields[v].n_steps; --j) g_order[j] = g_order[j - 1];
      g_order[j] = v;
    }
  int64_t acc = 0;
  for (int j = 0; j < n_fields; ++j) {
    g_stage.table[j] = g_table[g_order[j]];
    g_stage.prefix[j] = acc;
    acc += g_blocks[g_order[j]];
  }
  g_stage.prefix[n_fields] = acc;

  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(std::string("batch_run_noalloc: ") + hipGetErrorString(e), PYFTLE_EHIP);
  NoallocScratch *d = nullptr;
  rc = scratch_address(dev, &d);
  if (rc) return rc;
  const void *kernel = batch_kernel_for(shape.dim, opts->method, opts->time_mode, layout_pair(shape.layout),
                                        layout_cell(shape.layout), true);
  int grid = 0;
  std::string gerr;
  if (persistent_grid(kernel, shape.threads, opts->blocks_per_cu, shape.n_items, &grid, &gerr))
    return fail("batch_run_noalloc: launch shape: " + gerr, PYFTLE_EHIP);
  hipStream_t st = as_stream(stream);
  // a copy from pageable host memory returns once the source has been read, so
  // g_stage may be rewritten by the next call as soon as this one returns
  const size_t prefix_bytes = sizeof(int64_t) * (size_t)(n_fields + 1);
  e = hipMemcpyAsync(d->prefix, g_stage.prefix, prefix_bytes, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d->table, g_stage.table, sizeof(AdvectArgs) * (size_t)n_fields, hipMemcpyHostToDevice, st);
  if (e != hipSuccess)
    return fail(std::string("batch_run_noalloc: descriptor copy: ") + hipGetErrorString(e), PYFTLE_EHIP);
  reset_counter_kernel<<<1, 1, 0, st>>>(&d->counter);  // a kernel, not a memset
  rc = check_launch("batch_run_noalloc: counter reset");
  if (rc) return rc;
  const AdvectArgs *table = d->table;
  const int64_t *prefix = d->prefix;
  int nf = n_fields;
  int n = (int)shape.n_items;
  unsigned int *counter = &d->counter;
  void *args[] = {(void *)&table, (void *)&prefix, (void *)&nf, (void *)&n, (void *)&counter};
  e = hipLaunchKernel(kernel, dim3(grid), dim3(shape.threads), args, 0, st);
  if (e != hipSuccess) return fail(std::string("batch_run_noalloc: launch: ") + hipGetErrorString(e), PYFTLE_EHIP);
  return check_launch("batch_run_noalloc: launch");
}

}  // extern "C"
