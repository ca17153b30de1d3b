// Ragged batches (include/tsdx_hip.h, tsdx_hip_ragged_*): planning on the
// host, the autotuner, and the captured graph that a run replays.
//
// Plan:   sizes and kinds are read once; matrices are bucketed by (kind,
//         n <= NB), sorted by decreasing n inside a bucket and cut into wave
//         tasks; each register class gets its task list, longest tasks first.
// Tuner:  per class, every compiled variant (waves per workgroup, rows per
//         lane of the Cholesky buckets) factors a sample of the class's sizes
//         on synthetic matrices (the caller's data is never touched); the
//         fastest variant is kept, its persistent grid filling the GPU.
// Graph:  a counter reset kernel, then the class kernels on three forked
//         streams, joined; one hipGraphLaunch per run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/tsdx_hip.h"
#include "capi_internal.hpp"
#include "ragged_api.hpp"
#include "ragged_kernels.hpp"

namespace tsdx_hip {

__global__ void k_ragged_reset3(int* counters) {
  if (threadIdx.x < 3) counters[threadIdx.x] = 0;
}

namespace {

constexpr int kClasses = 3;

int nvariants(bool dbl, int cls) { return cls == 0 ? 3 : (cls == 1 ? 2 : (dbl ? 2 : 3)); }

// Type-erased access to the compiled (class, variant) launchers.
template <class T>
struct VariantOps {
  hipError_t (*launch)(const RaggedArgs<T>&, const RTask*, int, int*, int, hipStream_t);
  int (*blocks_per_cu)();
  int (*capacity)(int);
  int (*wpb)();
};
template <class T, int C, int V>
VariantOps<T> ops_of() {
  return {&launch_ragged<T, C, V>, &ragged_blocks_per_cu<T, C, V>, &ragged_capacity<T, C, V>, &ragged_wpb<T, C, V>};
}
template <class T>
VariantOps<T> variant_ops(int cls, int var);
template <>
VariantOps<float> variant_ops<float>(int cls, int var) {
  if (cls == 0) return var == 0 ? ops_of<float, 0, 0>() : var == 1 ? ops_of<float, 0, 1>() : ops_of<float, 0, 2>();
  if (cls == 1) return var == 0 ? ops_of<float, 1, 0>() : ops_of<float, 1, 1>();
  return var == 0 ? ops_of<float, 2, 0>() : var == 1 ? ops_of<float, 2, 1>() : ops_of<float, 2, 2>();
}
template <>
VariantOps<double> variant_ops<double>(int cls, int var) {
  if (cls == 0) return var == 0 ? ops_of<double, 0, 0>() : var == 1 ? ops_of<double, 0, 1>() : ops_of<double, 0, 2>();
  if (cls == 1) return var == 0 ? ops_of<double, 1, 0>() : ops_of<double, 1, 1>();
  return var == 0 ? ops_of<double, 2, 0>() : ops_of<double, 2, 3>();  // variant 3: 3 waves per SIMD
}

int bucket_code(int kind, int n) {
  if (kind == TSDX_HIP_KIND_CHOLESKY)
    return n <= 4 ? RC_CHOL4 : n <= 8 ? RC_CHOL8 : n <= 16 ? RC_CHOL16 : n <= 32 ? RC_CHOL32 : RC_CHOL64;
  return n <= 4 ? RC_LU4 : n <= 8 ? RC_LU8 : n <= 16 ? RC_LU16 : n <= kLuTileMaxN ? RC_LU32T : n <= 32 ? RC_LU32S : RC_LU64;
}

// Matrices (ids into the batch) of each bucket, sorted by decreasing n.
struct Buckets {
  std::vector<int> ids[RC_NCODES];
};
Buckets make_buckets(const std::vector<int>& size, const std::vector<int>& kind, const std::vector<int>& pick) {
  Buckets b;
  // counting sort by n (descending) keeps it linear for 10^6 matrices
  std::vector<int> byn[65];
  for (int m : pick) byn[size[m]].push_back(m);
  for (int n = 64; n >= 1; --n)
    for (int m : byn[n]) b.ids[bucket_code(kind[m], n)].push_back(m);
  return b;
}

// order (concatenated buckets of one class) and its wave tasks, longest first.
void class_tasks(const Buckets& b, int cls, int (*capacity)(int), const std::vector<int>& size, std::vector<int>& order,
                 std::vector<RTask>& tasks) {
  std::vector<std::pair<int64_t, RTask>> keyed;
  for (int code = 0; code < RC_NCODES; ++code) {
    if (rclass(code) != cls || b.ids[code].empty()) continue;
    const int cap = capacity(code);
    const std::vector<int>& ids = b.ids[code];
    for (size_t s = 0; s < ids.size(); s += (size_t)cap) {
      RTask t;
      t.code = code;
      t.first = (int)(order.size() + s);
      t.count = (int)std::min<size_t>((size_t)cap, ids.size() - s);
      t.nmax = size[ids[s]];
      const int64_t nm = t.nmax;
      const int64_t cost = nm * nm * nm * (code >= RC_LU4 ? 2 : 1) * (int64_t)t.count;
      keyed.push_back({cost, t});
    }
    order.insert(order.end(), ids.begin(), ids.end());
  }
  std::stable_sort(keyed.begin(), keyed.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
  tasks.clear();
  for (auto& k : keyed) tasks.push_back(k.second);
}

template <class V>
V* upload(const std::vector<V>& h) {
  if (h.empty()) return nullptr;
  V* d = nullptr;
  hip_check(hipMalloc(&d, h.size() * sizeof(V)), "ragged plan allocation");
  hip_check(hipMemcpy(d, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice), "ragged plan upload");
  return d;
}

}  // namespace

}  // namespace tsdx_hip

struct tsdx_hip_ragged_plan {
  int dbl = 0;
  int64_t count = 0;
  void* A = nullptr;
  const int64_t* off = nullptr;
  int* size = nullptr;  // plan-owned copy of the sizes read at creation
  int* piv = nullptr;
  const int64_t* poff = nullptr;
  int* info = nullptr;
  int* order = nullptr;
  tsdx_hip::RTask* tasks[3] = {nullptr, nullptr, nullptr};
  int ntasks[3] = {0, 0, 0};
  int var[3] = {0, 0, 0};
  int grid[3] = {0, 0, 0};
  int* counters = nullptr;
  hipStream_t side[3] = {nullptr, nullptr, nullptr};
  hipStream_t cap = nullptr;
  hipEvent_t fork = nullptr, join[3] = {nullptr, nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string report;
};

namespace tsdx_hip {
namespace {

void destroy_plan(tsdx_hip_ragged_plan* p) {
  if (!p) return;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  for (int c = 0; c < 3; ++c) {
    if (p->tasks[c]) (void)hipFree(p->tasks[c]);
    if (p->side[c]) (void)hipStreamDestroy(p->side[c]);
    if (p->join[c]) (void)hipEventDestroy(p->join[c]);
  }
  if (p->fork) (void)hipEventDestroy(p->fork);
  if (p->cap) (void)hipStreamDestroy(p->cap);
  if (p->order) (void)hipFree(p->order);
  if (p->size) (void)hipFree(p->size);
  if (p->counters) (void)hipFree(p->counters);
  delete p;
}

template <class T>
RaggedArgs<T> args_of(const tsdx_hip_ragged_plan* p) {
  RaggedArgs<T> a;
  a.A = static_cast<T*>(p->A);
  a.off = p->off;
  a.size = p->size;
  a.piv = p->piv;
  a.poff = p->poff;
  a.info = p->info;
  a.order = p->order;
  return a;
}

// Times every variant of class cls on synthetic matrices with the sizes and
// kinds of a sample of the class; returns the fastest variant.
template <class T>
int tune_class(int cls, const std::vector<int>& size, const std::vector<int>& kind, const std::vector<int>& members,
               int ncu, std::string& rep) {
  const bool dbl = sizeof(T) == 8;
  const int nv = nvariants(dbl, cls);
  if (nv == 1 || members.empty()) return 0;
  // sample: every k-th member, at most 32768 matrices
  const size_t want = 32768;
  const size_t step = members.size() > want ? members.size() / want : 1;
  std::vector<int> ssize, skind, pick;
  std::vector<int64_t> soff;
  int64_t total = 0;
  for (size_t s = 0; s < members.size() && ssize.size() < want; s += step) {
    const int m = members[s];
    pick.push_back((int)ssize.size());
    ssize.push_back(size[m]);
    skind.push_back(kind[m]);
    soff.push_back(total);
    total += (int64_t)size[m] * size[m] + 4;  // keep rows of every matrix 16-byte aligned when n allows
    total = (total + 3) & ~int64_t(3);
  }
  const int cnt = (int)ssize.size();
  std::vector<int64_t> spoff(cnt);
  int64_t ptot = 0;
  for (int i = 0; i < cnt; ++i) {
    spoff[i] = ptot;
    ptot += ssize[i];
  }
  T* dA = nullptr;
  int *dsize = upload(ssize), *dpiv = nullptr, *dinfo = nullptr, *dorder = nullptr, *dcnt = nullptr;
  int64_t *doff = upload(soff), *dpoff = upload(spoff);
  hip_check(hipMalloc(&dA, (size_t)total * sizeof(T)), "ragged tuner allocation");
  hip_check(hipMalloc(&dpiv, (size_t)ptot * sizeof(int)), "ragged tuner allocation");
  hip_check(hipMalloc(&dinfo, (size_t)cnt * sizeof(int)), "ragged tuner allocation");
  hip_check(hipMalloc(&dcnt, sizeof(int)), "ragged tuner allocation");
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  const Buckets b = make_buckets(ssize, skind, pick);
  int best = 0;
  float best_ms = 1e30f;
  for (int v = 0; v < nv; ++v) {
    const VariantOps<T> op = variant_ops<T>(cls, v);
    std::vector<int> order;
    std::vector<RTask> tasks;
    class_tasks(b, cls, op.capacity, ssize, order, tasks);
    if (dorder) hip_check(hipFree(dorder), "free");
    dorder = upload(order);
    RTask* dtasks = upload(tasks);
    RaggedArgs<T> a{dA, doff, dsize, dpiv, dpoff, dinfo, dorder};
    const int grid = std::max(1, op.blocks_per_cu() * ncu);
    float ms_min = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      hipLaunchKernelGGL(k_ragged_synth<T>, dim3((unsigned)cnt), dim3(256), 0, nullptr, dA, doff, dsize, cnt);
      hipLaunchKernelGGL(k_ragged_reset3, dim3(1), dim3(64), 0, nullptr, dcnt);
      hip_check(hipEventRecord(e0, nullptr), "event");
      hip_check(op.launch(a, dtasks, (int)tasks.size(), dcnt, grid, nullptr), "ragged tuner launch");
      hip_check(hipEventRecord(e1, nullptr), "event");
      hip_check(hipEventSynchronize(e1), "ragged tuner run");
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
      ms_min = std::min(ms_min, ms);
    }
    hip_check(hipFree(dtasks), "free");
    char line[160];
    snprintf(line, sizeof line, "\"class%d_variant%d_ms\": %.4f, ", cls, v, ms_min);
    rep += line;
    if (ms_min < best_ms) {
      best_ms = ms_min;
      best = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (void* q : {(void*)dA, (void*)dsize, (void*)dpiv, (void*)dinfo, (void*)dorder, (void*)dcnt, (void*)doff,
                  (void*)dpoff})
    if (q) (void)hipFree(q);
  return best;
}

template <class T>
tsdx_hip_ragged_plan* create_plan(T* A, const int64_t* off, const int* sizes, const int* kinds, int64_t count, int* piv,
                                  const int64_t* poff, int* info, int tune, hipStream_t s) {
  if (count < 0) throw InvalidArg("negative count");
  if (count > 0x7fffffff) throw InvalidArg("count above 2^31-1");
  if (count > 0 && (!A || !off || !sizes || !kinds)) throw InvalidArg("A, offsets, sizes and kinds must not be null");
  require_device();
  std::vector<int> size((size_t)count), kind((size_t)count);
  if (count > 0) {
    hip_check(hipStreamSynchronize(s), "ragged plan");
    hip_check(hipMemcpy(size.data(), sizes, (size_t)count * sizeof(int), hipMemcpyDeviceToHost), "ragged sizes");
    hip_check(hipMemcpy(kind.data(), kinds, (size_t)count * sizeof(int), hipMemcpyDeviceToHost), "ragged kinds");
  }
  bool any_lu = false;
  for (int64_t m = 0; m < count; ++m) {
    if (size[m] < 1 || size[m] > 64)
      throw InvalidArg("matrix " + std::to_string(m) + ": size n=" + std::to_string(size[m]) +
                       " outside supported range [1, 64]");
    if (kind[m] != TSDX_HIP_KIND_CHOLESKY && kind[m] != TSDX_HIP_KIND_LU)
      throw InvalidArg("matrix " + std::to_string(m) + ": unknown kind " + std::to_string(kind[m]));
    any_lu = any_lu || kind[m] == TSDX_HIP_KIND_LU;
  }
  if (any_lu && piv && !poff) throw InvalidArg("piv given without piv_offsets");

  auto* p = new tsdx_hip_ragged_plan();
  try {
    p->dbl = sizeof(T) == 8;
    p->count = count;
    p->A = A;
    p->off = off;
    // the kernels read the sizes the plan was built from, not the caller's
    // buffer: buckets, task bounds and register classes were fixed from them
    p->size = count > 0 ? upload(size) : nullptr;
    p->piv = piv;
    p->poff = poff;
    p->info = info;
    int dev = 0, ncu = 0;
    hip_check(hipGetDevice(&dev), "device");
    hip_check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "device attribute");
    std::vector<int> all((size_t)count);
    for (int64_t m = 0; m < count; ++m) all[m] = (int)m;
    const Buckets b = make_buckets(size, kind, all);
    std::vector<int> members[3];
    for (int code = 0; code < RC_NCODES; ++code)
      members[rclass(code)].insert(members[rclass(code)].end(), b.ids[code].begin(), b.ids[code].end());
    std::string rep = "{";
    for (int c = 0; c < kClasses; ++c) {
      p->var[c] = tune ? tune_class<T>(c, size, kind, members[c], ncu, rep) : 0;
    }
    // final order and tasks
    std::vector<int> order;
    for (int c = 0; c < kClasses; ++c) {
      const VariantOps<T> op = variant_ops<T>(c, p->var[c]);
      std::vector<RTask> tasks;
      std::vector<int> ord;
      class_tasks(b, c, op.capacity, size, ord, tasks);
      for (auto& t : tasks) t.first += (int)order.size();
      order.insert(order.end(), ord.begin(), ord.end());
      p->ntasks[c] = (int)tasks.size();
      p->tasks[c] = upload(tasks);
      p->grid[c] = std::max(1, std::min(p->ntasks[c], op.blocks_per_cu() * ncu));
      char line[200];
      snprintf(line, sizeof line,
               "\"class%d\": {\"matrices\": %zu, \"tasks\": %d, \"variant\": %d, \"waves_per_block\": %d, "
               "\"grid\": %d}, ",
               c, members[c].size(), p->ntasks[c], p->var[c], op.wpb(), p->grid[c]);
      rep += line;
    }
    p->order = upload(order);
    hip_check(hipMalloc(&p->counters, 4 * sizeof(int)), "ragged plan allocation");
    rep += "\"count\": " + std::to_string(count) + "}";
    p->report = rep;

    // capture: reset, fork, class kernels, join
    hip_check(hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking), "stream");
    hip_check(hipEventCreateWithFlags(&p->fork, hipEventDisableTiming), "event");
    for (int c = 0; c < kClasses; ++c) {
      hip_check(hipStreamCreateWithFlags(&p->side[c], hipStreamNonBlocking), "stream");
      hip_check(hipEventCreateWithFlags(&p->join[c], hipEventDisableTiming), "event");
    }
    const RaggedArgs<T> a = args_of<T>(p);
    hip_check(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal), "graph capture");
    hipLaunchKernelGGL(k_ragged_reset3, dim3(1), dim3(64), 0, p->cap, p->counters);
    hip_check(hipEventRecord(p->fork, p->cap), "graph capture");
    for (int c = 0; c < kClasses; ++c) {
      if (p->ntasks[c] == 0) continue;
      hip_check(hipStreamWaitEvent(p->side[c], p->fork, 0), "graph capture");
      hip_check(variant_ops<T>(c, p->var[c]).launch(a, p->tasks[c], p->ntasks[c], p->counters + c, p->grid[c], p->side[c]),
                "graph capture launch");
      hip_check(hipEventRecord(p->join[c], p->side[c]), "graph capture");
      hip_check(hipStreamWaitEvent(p->cap, p->join[c], 0), "graph capture");
    }
    hip_check(hipStreamEndCapture(p->cap, &p->graph), "graph capture");
    hip_check(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0), "graph instantiate");
  } catch (...) {
    destroy_plan(p);
    throw;
  }
  return p;
}

}  // namespace
}  // namespace tsdx_hip

using namespace tsdx_hip;

extern "C" {

int tsdx_hip_ragged_plan_create_f32(float* A, const int64_t* offsets, const int* sizes, const int* kinds, int64_t count,
                                    int* piv, const int64_t* piv_offsets, int* info, int tune,
                                    tsdx_hip_ragged_plan** plan, void* stream) {
  return guarded([&] {
    if (!plan) throw InvalidArg("plan must not be null");
    *plan = create_plan<float>(A, offsets, sizes, kinds, count, piv, piv_offsets, info, tune,
                               static_cast<hipStream_t>(stream));
  });
}
int tsdx_hip_ragged_plan_create_f64(double* A, const int64_t* offsets, const int* sizes, const int* kinds,
                                    int64_t count, int* piv, const int64_t* piv_offsets, int* info, int tune,
                                    tsdx_hip_ragged_plan** plan, void* stream) {
  return guarded([&] {
    if (!plan) throw InvalidArg("plan must not be null");
    *plan = create_plan<double>(A, offsets, sizes, kinds, count, piv, piv_offsets, info, tune,
                                static_cast<hipStream_t>(stream));
  });
}
int tsdx_hip_ragged_run(tsdx_hip_ragged_plan* plan, void* stream) {
  return guarded([&] {
    if (!plan) throw InvalidArg("plan must not be null");
    if (plan->count == 0) return;
    hip_check(hipGraphLaunch(plan->exec, static_cast<hipStream_t>(stream)), "ragged graph launch");
  });
}
const char* tsdx_hip_ragged_plan_report(const tsdx_hip_ragged_plan* plan) { return plan ? plan->report.c_str() : ""; }
int tsdx_hip_ragged_plan_destroy(tsdx_hip_ragged_plan* plan) {
  return guarded([&] { destroy_plan(plan); });
}

}  // extern "C"
