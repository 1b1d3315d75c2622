// Batched FTLE windows: one persistent launch over a work list of
// (field, block of seeds) items drawn from a device counter.  The per-item
// work is the cell-local FP32 window of advect_impl.hpp with its FTLE
// epilogue; each field brings its own grid, snapshot ring, seeds, step count,
// step size and time origin through a descriptor table in device memory.
#include <algorithm>
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

using BatchFn = void (*)(const AdvectArgs *, const int2 *, int, unsigned int *);

template <int DIM, int METHOD, int TMODE, bool PAIR>
static const void *kernel_ptr(int cell) {
  return cell ? reinterpret_cast<const void *>(&batch_kernel<DIM, METHOD, TMODE, PAIR, true>)
              : reinterpret_cast<const void *>(&batch_kernel<DIM, METHOD, TMODE, PAIR, false>);
}

template <int DIM, int METHOD>
static const void *pick_time(int tmode, int pair, int cell) {
  if (tmode == PYFTLE_TIME_LINEAR) return kernel_ptr<DIM, METHOD, PYFTLE_TIME_LINEAR, true>(cell);
  if (tmode == PYFTLE_TIME_FROZEN)
    return pair ? kernel_ptr<DIM, METHOD, PYFTLE_TIME_FROZEN, true>(cell)
                : kernel_ptr<DIM, METHOD, PYFTLE_TIME_FROZEN, false>(cell);
  return pair ? kernel_ptr<DIM, METHOD, PYFTLE_TIME_STEADY, true>(cell)
              : kernel_ptr<DIM, METHOD, PYFTLE_TIME_STEADY, false>(cell);
}

template <int DIM>
static const void *pick_method(int method, int tmode, int pair, int cell) {
  if (method == PYFTLE_RK4) return pick_time<DIM, PYFTLE_RK4>(tmode, pair, cell);
  if (method == PYFTLE_EULER) return pick_time<DIM, PYFTLE_EULER>(tmode, pair, cell);
  return pick_time<DIM, PYFTLE_AB2>(tmode, pair, cell);
}

static const void *batch_kernel_for(const pyftle_batch *b) {
  return b->dim == 2 ? pick_method<2>(b->method, b->tmode, b->pair, b->cell)
                     : pick_method<3>(b->method, b->tmode, b->pair, b->cell);
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

}  // namespace pyftle

using namespace pyftle;

extern "C" {

int pyftle_batch_create(const pyftle_batch_field *fields, int32_t n_fields,
                        const pyftle_batch_opts *opts, pyftle_batch **out) {
  if (!out) return fail("batch_create: out is NULL");
  *out = nullptr;
  if (!fields || n_fields <= 0) return fail("batch_create: no fields");
  if (!opts) return fail("batch_create: opts is NULL");
  if (opts->method != PYFTLE_EULER && opts->method != PYFTLE_AB2 && opts->method != PYFTLE_RK4)
    return fail("batch_create: unknown integrator");
  if (opts->time_mode != PYFTLE_TIME_STEADY && opts->time_mode != PYFTLE_TIME_FROZEN &&
      opts->time_mode != PYFTLE_TIME_LINEAR)
    return fail("batch_create: unknown time mode");
  if (opts->pos_dtype != PYFTLE_F32 && opts->pos_dtype != PYFTLE_F64)
    return fail("batch_create: bad positions dtype");
  if (opts->lane_order != PYFTLE_ORDER_SEED_MAJOR && opts->lane_order != PYFTLE_ORDER_NEIGHBOR_MAJOR)
    return fail("batch_create: unknown lane order");
  const int dim = fields[0].field.dim;
  const int layout = fields[0].field.layout;
  const int nb = 2 * dim;
  const int G = opts->groups_per_block > 0 ? opts->groups_per_block : 64;
  if (G * nb > 512) return fail("batch_create: groups_per_block * 2*dim must be <= 512");
  std::vector<AdvectArgs> table(n_fields);
  std::vector<int64_t> blocks(n_fields);
  int64_t n_items = 0;
  for (int i = 0; i < n_fields; ++i) {
    const pyftle_batch_field &bf = fields[i];
    int rc = make_field_dev(&bf.field, &table[i].f);
    if (rc) return fail("batch_create: field " + std::to_string(i) + ": " + pyftle_last_error(), rc);
    if (bf.field.dim != dim || bf.field.layout != layout)
      return fail("batch_create: all fields must share dim and record layout");
    if (bf.field.dtype != PYFTLE_F32) return fail("batch_create: batches run FP32 fields");
    if (bf.field.interp != PYFTLE_INTERP_GINTERP)
      return fail("batch_create: batches run GridInterp-mode fields (interp 0)");
    if (opts->time_mode == PYFTLE_TIME_LINEAR &&
        (!layout_pair(layout) || !(bf.field.dt_snap != 0.0)))
      return fail("batch_create: LINEAR time needs PAIR or CELL_PAIR fields with non-zero dt_snap");
    if (opts->time_mode == PYFTLE_TIME_FROZEN &&
        (bf.first_snapshot < 0 || (int64_t)bf.first_snapshot + bf.n_steps > bf.field.n_snapshots))
      return fail("batch_create: FROZEN steps run past the last snapshot of field " + std::to_string(i));
    if (opts->time_mode == PYFTLE_TIME_STEADY &&
        (bf.first_snapshot < 0 || bf.first_snapshot >= bf.field.n_snapshots))
      return fail("batch_create: STEADY snapshot out of range in field " + std::to_string(i));
    if (bf.n_groups < 0 || bf.n_steps < 0) return fail("batch_create: negative size");
    if (bf.n_groups > 0 && (!bf.positions_in || !bf.ftle_out))
      return fail("batch_create: NULL positions or output in field " + std::to_string(i));
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
  if (n_items >= INT32_MAX) return fail("batch_create: too many work items");
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
  b->dim = dim;
  b->method = opts->method;
  b->tmode = opts->time_mode;
  b->pair = layout_pair(layout);
  b->cell = layout_cell(layout);
  b->threads = ((G * nb + 63) / 64) * 64;
  b->n_items = n_items;
  b->use_graph = opts->use_graph != 0;
  auto hip_fail = [&](const char *what, hipError_t e) {
    release(b);
    return fail(std::string("batch_create: ") + what + ": " + hipGetErrorString(e), PYFTLE_EHIP);
  };
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail("hipGetDevice", e);
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return hip_fail("device attribute", e);
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel_for(b), b->threads, 0);
  if (e != hipSuccess) return hip_fail("occupancy", e);
  per_cu = std::max(1, per_cu);
  if (opts->blocks_per_cu > 0) per_cu = std::min(per_cu, (int)opts->blocks_per_cu);
  b->grid = std::max(1, (int)std::min<int64_t>((int64_t)cus * per_cu, std::max<int64_t>(n_items, 1)));
  e = hipMalloc(&b->d_table, sizeof(AdvectArgs) * n_fields);
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

}  // extern "C"
