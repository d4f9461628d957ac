// batch.hip -- the batch runner's persistent work-queue kernels (no
// counterpart in awkward-cpp).
//
// A batch is many independent jagged arrays of one dtype (zero-based int64
// offsets, contiguous content).  Each array is cut into tiles of lists; the
// tiles of all arrays are numbered consecutively (tile_begin per array,
// computed on the host from the list counts) and a fixed grid of persistent
// workgroups pulls tile numbers from an atomic counter until the queue is
// empty -- so uneven arrays balance themselves across the chip and the whole
// batch is one launch per dtype and op, capturable in a hipGraph.
//
//   awkward_hip_batch_reduce_<T>: per list, sum (awkward_reduce_sum's output
//     type), min (awkward_reduce_min with its identity) and the local argmin
//     with the empty-list mask (awkward_reduce_argmin + adjust_starts +
//     reduce_mask), all from one pass over the elements, with one of the ten
//     reduce tiles; awkward_hip_batch_reduce_<T>_chosen enqueues all ten,
//     predicated on a tile chosen on the device (learned.hip);
//   awkward_hip_batch_argsort_<T>: per list, the stable argsort
//     (awkward_argsort), local indices.
//
// The counter lives in device memory that the caller clears before each
// launch (awkward_hip_batch_reset: a kernel, so the sequence can be graph
// captured).  Every workgroup leaves the loop when the counter passes the
// last tile, so the grid always drains.

#include "reduce.h"
#include "sort.h"

namespace akhip {
namespace {

struct BatchReduceItem {
  const void* data;
  const int64_t* offsets;
  void* out_sum;
  void* out_min;
  int64_t* out_argmin;
  int8_t* out_mask;
  int64_t n;
  int64_t tile_begin;
};

struct BatchSortItem {
  const void* data;
  const int64_t* offsets;
  int64_t* out;
  int64_t n;         // lists
  int64_t length;    // elements
  int64_t tile_begin;
  int64_t elem_begin;  // this array's region of the merge buffers
};

template <typename T, typename S>
struct BatchEpilogue {
  S* sum;
  T* mn;
  int64_t* arg;
  int8_t* mask;
  __device__ void store(int64_t bin, MultiOut<T, S> v, int64_t gb, int64_t ge) const {
    sum[bin] = v.sum;
    mn[bin] = v.min;
    arg[bin] = v.arg >= 0 ? v.arg - gb : -1;
    mask[bin] = (ge > gb) ? 0 : 1;
  }
};

// Per-item data and epilogue for the long-list queue's kernels.
template <typename T, typename S>
struct BatchSrc {
  const BatchReduceItem* items;
  __device__ const T* data(int i) const { return (const T*)items[i].data; }
  __device__ BatchEpilogue<T, S> epi(int i) const {
    const BatchReduceItem& it = items[i];
    return BatchEpilogue<T, S>{(S*)it.out_sum, (T*)it.out_min, it.out_argmin, it.out_mask};
  }
};

// item owning global tile `tile`: the last one with tile_begin <= tile
template <typename Item>
__device__ inline int find_item(const Item* items, int nitems, int64_t tile) {
  int lo = 0, hi = nitems;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (items[mid].tile_begin <= tile) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ void reset_counter(unsigned long long* counter) { *counter = 0ull; }

template <typename T, typename S, int BLOCK, int LPT, int CAP>
__global__ __launch_bounds__(BLOCK) void batch_reduce_kernel(const BatchReduceItem* __restrict__ items, int nitems,
                                                             int64_t total_tiles, unsigned long long* counter,
                                                             T identity, const int32_t* __restrict__ choice,
                                                             int self, ReduceQueue q) {
  using Op = SumMinArgminOp<T, S>;
  __shared__ ReduceSmem<Op, BLOCK, LPT, CAP> sm;
  __shared__ int64_t s_tile;
  __shared__ int s_item;
  // a launch predicated on a device-side choice: every workgroup leaves at once
  // unless this variant was chosen
  if (choice != nullptr && choice[0] != self) return;
  const Op op{identity};
  for (;;) {
    if (threadIdx.x == 0) {
      const int64_t t = (int64_t)atomicAdd(counter, 1ull);
      s_tile = t;
      s_item = t < total_tiles ? find_item(items, nitems, t) : 0;
    }
    __syncthreads();
    const int64_t tile = s_tile;
    if (tile >= total_tiles) break;
    const int item = s_item;
    const BatchReduceItem it = items[item];
    __syncthreads();  // s_tile / s_item are rewritten by the next claim
    const BatchEpilogue<T, S> epi{(S*)it.out_sum, (T*)it.out_min, it.out_argmin, it.out_mask};
    reduce_tile<Op, BLOCK, LPT, CAP>(op, epi, (const T*)it.data, it.offsets, it.n,
                                     (tile - it.tile_begin) * (int64_t)(BLOCK * LPT), sm, q, item);
  }
}

template <typename T>
__global__ __launch_bounds__(sortimpl::kBlock) void batch_argsort_kernel(const BatchSortItem* __restrict__ items,
                                                                         int nitems, int64_t total_tiles,
                                                                         unsigned long long* counter, bool asc,
                                                                         sortimpl::LongQueue q) {
  using namespace sortimpl;
  __shared__ int64_t s_tile;
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) {
      const int64_t t = (int64_t)atomicAdd(counter, 1ull);
      s_tile = t;
      s_item = t < total_tiles ? find_item(items, nitems, t) : 0;
    }
    __syncthreads();
    const int64_t tile = s_tile;
    if (tile >= total_tiles) break;
    const BatchSortItem it = items[s_item];
    __syncthreads();
    sort_tile<T, int64_t, true>(it.out, (const T*)it.data, it.length, it.offsets, it.n,
                                (tile - it.tile_begin) * (int64_t)kBlock, asc, q, it.elem_begin);
  }
}

template <typename T, typename S, int BLOCK, int LPT, int CAP>
Error launch_batch_reduce(const BatchReduceItem* items, int nitems, int64_t total_tiles, unsigned long long* counter,
                          T identity, const int32_t* choice, int self, const ReduceQueue& q) {
  // persistent grid: a few workgroups per CU, never more than the tiles
  int64_t grid = (int64_t)cu_count() * (BLOCK <= 64 ? 16 : BLOCK <= 128 ? 8 : 4);
  if (grid > total_tiles) grid = total_tiles;
  if (grid < 1) grid = 1;
  batch_reduce_kernel<T, S, BLOCK, LPT, CAP><<<(unsigned)grid, BLOCK, 0, stream()>>>(
      items, nitems, total_tiles, counter, identity, choice, self, q);
  AKHIP_LAUNCH_CHECK("batch_reduce_kernel");
  return success();
}

// The ten tiles of launch_reduce (threads x lists per thread x LDS bytes).
template <typename T, typename S>
Error launch_batch_variant(int variant, const BatchReduceItem* it, int nitems, int64_t tiles,
                           unsigned long long* counter, T identity, const int32_t* choice, const ReduceQueue& q) {
  switch (variant) {
    case 1: return launch_batch_reduce<T, S, 256, 4, 8192>(it, nitems, tiles, counter, identity, choice, 1, q);
    case 2: return launch_batch_reduce<T, S, 256, 2, 16384>(it, nitems, tiles, counter, identity, choice, 2, q);
    case 3: return launch_batch_reduce<T, S, 128, 1, 16384>(it, nitems, tiles, counter, identity, choice, 3, q);
    case 4: return launch_batch_reduce<T, S, 256, 1, 65536>(it, nitems, tiles, counter, identity, choice, 4, q);
    case 5: return launch_batch_reduce<T, S, 256, 8, 8192>(it, nitems, tiles, counter, identity, choice, 5, q);
    case 6: return launch_batch_reduce<T, S, 64, 1, 6144>(it, nitems, tiles, counter, identity, choice, 6, q);
    case 7: return launch_batch_reduce<T, S, 128, 2, 16384>(it, nitems, tiles, counter, identity, choice, 7, q);
    case 8: return launch_batch_reduce<T, S, 64, 4, 8192>(it, nitems, tiles, counter, identity, choice, 8, q);
    case 9: return launch_batch_reduce<T, S, 128, 1, 8192>(it, nitems, tiles, counter, identity, choice, 9, q);
    default: return launch_batch_reduce<T, S, 256, 1, 32768>(it, nitems, tiles, counter, identity, choice, 0, q);
  }
}

// The long-list queue of one batch reduction, from the queue pool: the
// count, the chunk table, then one partial per chunk.
template <typename T, typename S>
bool batch_queue(int64_t total_elements, ReduceQueue* q) {
  using Part = typename SumMinArgminOp<T, S>::Part;
  const int64_t cap = reduce_queue_capacity(total_elements > 0 ? total_elements : 1);
  const size_t cb = ((size_t)cap * sizeof(ReduceChunk) + 255) / 256 * 256;
  const size_t pb = ((size_t)cap * sizeof(Part) + 255) / 256 * 256;
  char* m = (char*)queue_pool(256 + cb + pb);
  if (!m) return false;
  *q = ReduceQueue{(ReduceChunk*)(m + 256), (void*)(m + 256 + cb), (unsigned long long*)m, cap};
  return true;
}

// tiles over all the batch's lists (variant `variant`, or all ten predicated
// on `choice`), then the queued long lists across the chip
template <typename T, typename S>
Error batch_reduce(const BatchReduceItem* const* tables, const int64_t* tiles, int nitems,
                   unsigned long long* counter, T identity, int variant, const int32_t* choice,
                   int64_t total_elements) {
  ReduceQueue q;
  if (!batch_queue<T, S>(total_elements, &q))
    return failure("out of device memory for the batch reduction queue", kSliceNone, kSliceNone,
                   "awkward_amd/csrc/batch.hip");
  hipStream_t st = stream();
  reduce_queue_reset<><<<1, 1, 0, st>>>(q.count);
  for (int v = 0; v < 10; v++) {
    const int k = choice ? v : variant;
    const int64_t nt = choice ? tiles[v] : tiles[0];
    if (nt > 0) {
      Error e = launch_batch_variant<T, S>(k, choice ? tables[v] : tables[0], nitems, nt, counter, identity, choice, q);
      if (e.str) return e;
    }
    if (!choice) break;
  }
  const SumMinArgminOp<T, S> op{identity};
  const BatchSrc<T, S> src{tables[0]};  // every table lists the same items in the same order
  int64_t grid = (int64_t)cu_count() * 8;
  reduce_chunks_kernel<SumMinArgminOp<T, S>, BatchSrc<T, S>><<<(unsigned)grid, 256, 0, st>>>(op, src, q);
  AKHIP_LAUNCH_CHECK("reduce_chunks_kernel");
  reduce_finalize_kernel<SumMinArgminOp<T, S>, BatchSrc<T, S>><<<(unsigned)cu_count(), 256, 0, st>>>(op, src, q);
  AKHIP_LAUNCH_CHECK("reduce_finalize_kernel");
  return success();
}

// scratch: the long-segment queue and two (key, index) merge buffers over
// the batch's elements, from the pool.  Segments longer than 4096 elements
// are sorted by every workgroup after the tile pass (sort.h).
template <typename T>
Error batch_argsort(const BatchSortItem* items, int32_t nitems, int64_t total_tiles, unsigned long long* counter,
                    bool ascending, int64_t total_elements) {
  using namespace sortimpl;
  if (nitems <= 0 || total_tiles <= 0) return success();
  const int64_t ne = total_elements > 0 ? total_elements : 1;
  const size_t kb = ((size_t)ne * 8 + 255) / 256 * 256;
  const size_t ib = ((size_t)ne * 4 + 255) / 256 * 256;
  const int64_t capacity = chunk_queue_capacity(ne);
  const int64_t small_capacity = small_queue_capacity(ne);
  const size_t qb = ((size_t)capacity * sizeof(LongChunk) + 255) / 256 * 256;
  const size_t sb = ((size_t)small_capacity * sizeof(LongChunk) + 255) / 256 * 256;
  const size_t cb = corank_bytes(capacity);
  char* s = (char*)scratch(256 + qb + sb + cb + 2 * (kb + ib));
  if (!s) return failure("out of device memory for batch sort scratch", kSliceNone, kSliceNone,
                         "awkward_amd/csrc/batch.hip");
  LongQueue q{(LongChunk*)(s + 256), (unsigned long long*)s, (unsigned long long*)(s + 8), capacity,
              (LongChunk*)(s + 256 + qb), (unsigned long long*)(s + 16), small_capacity,
              (int64_t*)(s + 256 + qb + sb)};
  char* b = s + 256 + qb + sb + cb;
  hipStream_t st = stream();
  long_queue_reset<><<<1, 1, 0, st>>>(q.count, q.maxlen, q.small_count);
  int64_t grid = (int64_t)cu_count() * 4;
  if (grid > total_tiles) grid = total_tiles;
  batch_argsort_kernel<T><<<(unsigned)grid, kBlock, 0, st>>>(items, nitems, total_tiles, counter, ascending, q);
  AKHIP_LAUNCH_CHECK("batch_argsort_kernel");
  const hipError_t e = launch_long_sort<T, true>(q, ne, ascending, (uint64_t*)b, (uint32_t*)(b + 2 * kb),
                                                 (uint64_t*)(b + kb), (uint32_t*)(b + 2 * kb + ib), st);
  if (e != hipSuccess) return hip_failure(e, "batch long segment sort");
  return success();
}

}  // namespace
}  // namespace akhip

using namespace akhip;

extern "C" {

// Lists per tile of each batch-reduce variant (the host numbers the tiles).
int64_t awkward_hip_batch_reduce_tile_lists(int variant) {
  static const int64_t lists[10] = {256, 1024, 512, 128, 256, 2048, 64, 256, 256, 128};
  return (variant >= 0 && variant < 10) ? lists[variant] : lists[0];
}

int64_t awkward_hip_batch_sort_tile_lists(void) { return sortimpl::kBlock; }

Error awkward_hip_batch_reset(unsigned long long* counter) {
  reset_counter<<<1, 1, 0, stream()>>>(counter);
  AKHIP_LAUNCH_CHECK("reset_counter");
  return success();
}

// tables / total_tiles (host arrays of 10): the item table and tile count
// built for each variant's lists per tile; choice (device): the variant to
// run (awkward_hip_dispatch_choose's or _infer's output).  All ten launches are enqueued,
// each leaves at once unless chosen, so no host synchronisation is needed.
// total_elements: the batch's element count (sizes the long-list queue).
#define BATCH_REDUCE(TN, T, S)                                                                                   \
  Error awkward_hip_batch_reduce_##TN(const void* items, int32_t nitems, int64_t total_tiles,                   \
                                      unsigned long long* counter, T identity, int32_t variant,                \
                                      int64_t total_elements) {                                                \
    if (nitems <= 0 || total_tiles <= 0) return success();                                                     \
    const BatchReduceItem* t = (const BatchReduceItem*)items;                                                  \
    return batch_reduce<T, S>(&t, &total_tiles, nitems, counter, identity, variant, nullptr, total_elements);  \
  }                                                                                                            \
  Error awkward_hip_batch_reduce_##TN##_chosen(const void* const* tables, const int64_t* total_tiles,          \
                                               int32_t nitems, unsigned long long* counter, T identity,        \
                                               const int32_t* choice, int64_t total_elements) {                \
    if (nitems <= 0) return success();                                                                         \
    return batch_reduce<T, S>((const BatchReduceItem* const*)tables, total_tiles, nitems, counter, identity,   \
                              0, choice, total_elements);                                                      \
  }
BATCH_REDUCE(int32, int32_t, int64_t)
BATCH_REDUCE(int64, int64_t, int64_t)
BATCH_REDUCE(float32, float, float)
BATCH_REDUCE(float64, double, double)
#undef BATCH_REDUCE

#define BATCH_SORT(TN, T)                                                                                      \
  Error awkward_hip_batch_argsort_##TN(const void* items, int32_t nitems, int64_t total_tiles,                \
                                       unsigned long long* counter, bool ascending, int64_t total_elements) { \
    return batch_argsort<T>((const BatchSortItem*)items, nitems, total_tiles, counter, ascending,            \
                            total_elements);                                                                 \
  }
BATCH_SORT(int32, int32_t)
BATCH_SORT(int64, int64_t)
BATCH_SORT(float32, float)
BATCH_SORT(float64, double)
#undef BATCH_SORT

}  // extern "C"
