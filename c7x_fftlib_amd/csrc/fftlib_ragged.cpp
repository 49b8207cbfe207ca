// Ragged batches: many complex FFTs of different lengths in one call, run as
// one HIP graph launch, with per-length kernel parameters chosen by timing
// (the autotuner) and kept in a process-wide cache that can be saved to and
// loaded from a text file.
//
// The reference has no ragged entry point (its batched kernels take one
// length per call, src/fft_c7x/FFTLIB_fft1dBatched_i32fc_c32fc_o32fc/
// FFTLIB_fft1dBatched_i32fc_c32fc_o32fc.h:92); a reference caller issues one
// _kernel call per length. Each transform here has the reference's C2C
// semantics: forward as FFTLIB_fft1dBatched_i32fc_c32fc_o32fc, inverse as
// FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc (conj, x(1.0f/N), conj), with the
// reference's twiddle tables.
//
// Two forms of plan, chosen per plan (FFTLIB_GPU_raggedSetForm) or by the
// autotuner:
//
// Persistent form (FFTLIB_RAGGED_PERSISTENT): one launch of the
// persistent ragged kernel (fft_ragged_class.hip) per direction runs every
// transform from a work queue (fftlib_ragged_queue.h). Lengths up to 4096 are
// plain tasks, one queue unit per group; each transform of a longer length
// n = n1 * m2 is one unit run by one workgroup: its column pass into its own
// scratch region, a barrier, its row pass.
//
// Graph form (FFTLIB_RAGGED_GRAPH, the default): transforms are grouped by
// length. Lengths up to 16384 become tasks of the multi-length kernel of
// their octave (fft_fixed_class.h). A longer length
// n = n1 * m2 runs in two passes through its own region of a plan-owned
// scratch buffer: a column task (n1-point DFTs of its m2 columns, twiddled)
// in the class kernel of n1, then a row task (m2-point DFTs of its n1 rows,
// written in natural order) in the class kernel of m2. One direction is
// therefore two phases: phase 0 = the lengths up to 16384 and the column
// tasks, phase 1 = the row tasks, each at most one class-kernel launch per
// octave spread over several capture streams, with a join between the
// phases (fftlib_ragged_schedule.h). Each class kernel walks a persistent
// list of (task, group) items, largest tasks first, so the whole device
// stays busy with many short transforms of different lengths at once.
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "fft_fixed_class.h"
#include "fft_fixed_impl.h"
#include "fft_kernels.h"
#include "fftlib_gpu.h"
#include "fftlib_ragged_queue.h"
#include "fftlib_ragged_schedule.h"
#include "fftlib_runtime.h"

using namespace fftlib;

struct FFTLIB_GPU_ragged_s {
   struct Group {
      uint32_t          n        = 0;
      uint32_t          count    = 0;
      bool              twoPass  = false;   // longer than the class kernels: two passes
      TwoPassPlan       plan{};
      uint64_t         *dOff     = nullptr; // channel offsets (complex elements)
      float            *dTw      = nullptr; // twiddle table of length n (one-pass lengths)
      uint64_t          scrOff   = 0;       // two-pass: complex offset of its scratch region
      double            work     = 0.0;     // n * log2(n) * count, for ordering
      // persistent form: lengths above kRaggedClassMaxN run two passes
      bool              pTwoPass = false;
      TwoPassPlan       pPlan{};
      uint64_t          pScrOff  = 0;
   };
   std::vector<Group> groups;
   struct Class {
      int                     cls      = -1;
      int                     phase    = 0;
      bool                    columns  = false;   // phase 0: holds column tasks of the longer lengths
      int                     chain    = -1;      // phase 1: octave of the column tasks of its transforms
      uint32_t                numItems = 0;
      fixedk::ClassTask      *dTasks   = nullptr;
      uint32_t               *dItems   = nullptr;
      uint32_t                grid     = 0;   // workgroups of the launch (0: the class default)
      double                  work     = 0.0;
   };
   std::vector<Class> classes;   // phase 0 first, then phase 1; by work within a phase
   // persistent form
   int                form        = FFTLIB_RAGGED_GRAPH;   // the faster form on the ragged 10000 workload
   bool               formSet     = false;   // chosen by FFTLIB_GPU_raggedSetForm (the tuner keeps it)
   bool                pAvailable = true;   // every length fits the persistent kernel (factors up to 4096)
   fixedk::ClassTask  *pTasks     = nullptr;
   fixedk::RaggedUnit *pUnits     = nullptr;
   uint32_t           *pCounters  = nullptr;
   uint32_t            pNumItems  = 0;   // units
   uint32_t           pGrid       = 0;       // workgroups of the persistent launch (0: resident)
   uint64_t           totalPoints = 0;
   float2            *scratch     = nullptr;
   size_t             scratchSize = 0;
   hipStream_t        capture     = nullptr;
   std::vector<hipStream_t> lanes;
   std::vector<hipEvent_t>  events;
   hipGraph_t         graph       = nullptr;
   hipGraphExec_t     exec        = nullptr;
   const void        *gx = nullptr, *gy = nullptr, *gz = nullptr;
   int                gdir        = -1;
   std::mutex         mu;
};

namespace {

constexpr uint32_t kRaggedMinN  = 16;
constexpr uint32_t kRaggedMaxN  = 1u << 20;
constexpr int      kCaptureLanes = 8;

bool smooth235(uint32_t n)
{
   for (uint32_t p : {2u, 3u, 5u})
      while (n % p == 0u) n /= p;
   return n == 1u;
}

// ---- tuning cache -------------------------------------------------------------
// Key: the class-kernel launch (kClassKeyBase + 100 * phase + octave index,
// plus 1000 * (1 + octave of the column tasks) for row launches) and a
// power-of-two bucket of its item count; value: the launch's number of
// workgroups (0 = the class kernels' default) and the time of a forward
// direction of the plan it was tuned in, with that grid.
constexpr uint32_t kClassKeyBase = 100000u;
uint32_t class_key(const FFTLIB_GPU_ragged_s::Class &c)
{
   return kClassKeyBase + 100u * (uint32_t) c.phase + (uint32_t) c.cls + (c.chain >= 0 ? 1000u * (uint32_t) (c.chain + 1) : 0u);
}
struct TuneKey {
   uint32_t n, bucket;
   bool     operator<(const TuneKey &o) const { return n != o.n ? n < o.n : bucket < o.bucket; }
};
struct TuneVal {
   uint32_t grid;
   float    us;
};
std::mutex                     g_tuneMu;
std::map<TuneKey, TuneVal>     g_tune;

uint32_t bucket_of(uint32_t count)
{
   uint32_t b = 1;
   while (b < count) b <<= 1;
   return b;
}

bool tuned(uint32_t n, uint32_t count, uint32_t *grid)
{
   std::lock_guard<std::mutex> lock(g_tuneMu);
   auto it = g_tune.find(TuneKey{n, bucket_of(count)});
   if (it == g_tune.end()) return false;
   *grid = it->second.grid;
   return true;
}

hipError_t launch_class(const FFTLIB_GPU_ragged_s &h, const FFTLIB_GPU_ragged_s::Class &c, const float *in, float *out,
                        int mode, uint32_t grid, hipStream_t s)
{
   fixedk::ClassArgs a{};
   a.x = (const float2 *) in, a.y = (float2 *) out, a.scratch = h.scratch, a.mode = mode;
   a.numItems = c.numItems, a.tasks = c.dTasks, a.items = c.dItems;
   return fixed_class_launch(c.cls, a, grid, s);
}

hipError_t launch_persistent(const FFTLIB_GPU_ragged_s &h, const float *in, float *out, int mode, uint32_t grid,
                             hipStream_t s)
{
   fixedk::ClassArgs a{};
   a.x = (const float2 *) in, a.y = (float2 *) out, a.scratch = h.scratch, a.mode = mode;
   a.numItems = h.pNumItems, a.tasks = h.pTasks, a.units = h.pUnits, a.counters = h.pCounters;
   return ragged_class_launch(a, grid, s);
}

FFTLIB_STATUS fail(const char *what, hipError_t e)
{
   set_last_error(what, e);
   return FFTLIB_ERR_FAILURE;
}

void release(FFTLIB_GPU_ragged_s *h)
{
   if (h->exec) (void) hipGraphExecDestroy(h->exec);
   if (h->graph) (void) hipGraphDestroy(h->graph);
   for (auto &g : h->groups) {
      if (g.dOff) (void) hipFree(g.dOff);
      if (g.dTw) (void) hipFree(g.dTw);
   }
   for (auto &c : h->classes) {
      if (c.dTasks) (void) hipFree(c.dTasks);
      if (c.dItems) (void) hipFree(c.dItems);
   }
   if (h->pTasks) (void) hipFree(h->pTasks);
   if (h->pUnits) (void) hipFree(h->pUnits);
   if (h->pCounters) (void) hipFree(h->pCounters);
   if (h->scratch) (void) hipFree(h->scratch);
   for (hipEvent_t e : h->events) (void) hipEventDestroy(e);
   for (hipStream_t s : h->lanes) (void) hipStreamDestroy(s);
   if (h->capture) (void) hipStreamDestroy(h->capture);
}

// Kind of each class launch, in the plan's order, and for each row launch
// the index of the column launch holding its transforms' column tasks.
std::vector<ragged::Kind> kinds_of(const FFTLIB_GPU_ragged_s *h)
{
   std::vector<ragged::Kind> k;
   for (const auto &c : h->classes)
      k.push_back(c.phase == 1 ? ragged::Kind::Row : c.columns ? ragged::Kind::Column : ragged::Kind::Plain);
   return k;
}
std::vector<int> links_of(const FFTLIB_GPU_ragged_s *h)
{
   std::vector<int> l(h->classes.size(), -1);
   for (size_t i = 0; i < h->classes.size(); ++i) {
      if (h->classes[i].phase != 1) continue;
      for (size_t j = 0; j < h->classes.size(); ++j)
         if (h->classes[j].phase == 0 && h->classes[j].columns && h->classes[j].cls == h->classes[i].chain)
            l[i] = (int) j;
   }
   return l;
}

// One direction's buffers and mode.
struct Direction {
   const float *in;
   float       *out;
   int          mode;
};

// Enqueues `count` directions by performing the steps of their schedule
// (fftlib_ragged_schedule.h) on the origin stream and the lanes. The pool
// holds ragged::events_needed of the plan's two-direction schedule; the index
// check turns a schedule that names more events into an error instead of a
// read past the pool.
hipError_t enqueue(FFTLIB_GPU_ragged_s *h, const Direction *dirs, int count, hipStream_t origin)
{
   hipError_t e = hipSuccess;
   if (h->form == FFTLIB_RAGGED_PERSISTENT) {   // one launch per direction, in stream order
      for (int d = 0; d < count && e == hipSuccess; ++d)
         e = launch_persistent(*h, dirs[d].in, dirs[d].out, dirs[d].mode, h->pGrid, origin);
      return e;
   }
   for (const ragged::Step &st : ragged::schedule(kinds_of(h), links_of(h), (int) h->lanes.size(), count)) {
      const hipStream_t s = st.stream == ragged::kOrigin ? origin : h->lanes[(size_t) st.stream];
      if (st.op == ragged::Op::Launch) {
         const auto      &c = h->classes[(size_t) st.cls];
         const Direction &d = dirs[st.dir];
         e                  = launch_class(*h, c, d.in, d.out, d.mode, c.grid, s);
      }
      else if (st.ev >= h->events.size()) e = hipErrorInvalidValue;
      else if (st.op == ragged::Op::Record) e = hipEventRecord(h->events[st.ev], s);
      else e = hipStreamWaitEvent(s, h->events[st.ev], 0);
      if (e != hipSuccess) return e;
   }
   return hipSuccess;
}

hipError_t capture_graph(FFTLIB_GPU_ragged_s *h, const float *x, float *y, float *z, int dir)
{
   if (h->exec) (void) hipGraphExecDestroy(h->exec), h->exec = nullptr;
   if (h->graph) (void) hipGraphDestroy(h->graph), h->graph = nullptr;
   hipError_t e = hipStreamBeginCapture(h->capture, hipStreamCaptureModeThreadLocal);
   if (e != hipSuccess) return e;
   const Direction dirs[2] = {{x, y, dir == 1 ? C2C_INVERSE_PRESCALE : C2C_FORWARD}, {y, z, C2C_INVERSE_PRESCALE}};
   const hipError_t le     = enqueue(h, dirs, dir == 2 ? 2 : 1, h->capture);
   hipGraph_t graph = nullptr;
   e                = hipStreamEndCapture(h->capture, &graph);
   if (le != hipSuccess) {
      if (graph) (void) hipGraphDestroy(graph);
      return le;
   }
   if (e != hipSuccess) return e;
   h->graph = graph;
   if ((e = hipGraphInstantiate(&h->exec, graph, nullptr, nullptr, 0)) != hipSuccess) return e;
   h->gx = x, h->gy = y, h->gz = z, h->gdir = dir;
   return hipSuccess;
}

// Makes the plan's capture stream wait for the work already enqueued on the
// caller's stream (the library stream set with FFTLIB_GPU_setStream).
hipError_t order_after_caller(FFTLIB_GPU_ragged_s *h)
{
   hipEvent_t ev;
   hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
   if (e != hipSuccess) return e;
   if ((e = hipEventRecord(ev, current_stream())) == hipSuccess) e = hipStreamWaitEvent(h->capture, ev, 0);
   (void) hipEventDestroy(ev);
   return e;
}

// Class-kernel tasks and items of the plan: one task per length up to 16384
// (phase 0) and a column task (phase 0) plus a row task (phase 1) per longer
// length, grouped by phase and by the octave of the task length; items
// (task << 20 | group) largest task first.
hipError_t build_classes(FFTLIB_GPU_ragged_s *h)
{
   // (phase, class, chain): row tasks are grouped by the octave of their
   // transforms' column tasks (chain), -1 for phase 0
   std::map<std::tuple<int, int, int>, std::vector<fixedk::ClassTask>> tasks;
   std::map<std::tuple<int, int, int>, double>                         work;
   for (const auto &g : h->groups) {
      std::vector<std::pair<int, fixedk::ClassTask>> ks;
      int                                             chain = -1;
      if (!g.twoPass) {
         fixedk::ClassTask k{};
         k.n = g.n, k.numChannels = g.count, k.layout = fixedk::LAYOUT_PLAIN, k.tw = (const float2 *) g.dTw;
         k.inOff = g.dOff, k.outOff = g.dOff;
         k.numGroups = (k.numChannels + class_group_size(k.n) - 1u) / class_group_size(k.n);
         ks.push_back({0, k});
      }
      else {
         ks.push_back({0, two_pass_task(g.plan, fixedk::LAYOUT_COLUMNS, g.count, g.dOff, g.scrOff)});
         ks.push_back({1, two_pass_task(g.plan, fixedk::LAYOUT_ROWS, g.count, g.dOff, g.scrOff)});
         chain = fixed_class_of(ks[0].second.n);
      }
      for (const auto &pk : ks) {
         const int cls = fixed_class_of(pk.second.n);
         if (cls < 0) return hipErrorInvalidValue;
         const auto key = std::make_tuple(pk.first, cls, pk.first == 1 ? chain : -1);
         tasks[key].push_back(pk.second);
         work[key] += g.work / (double) ks.size();
      }
   }
   for (auto &kv : tasks) {
      FFTLIB_GPU_ragged_s::Class c;
      c.phase = std::get<0>(kv.first);
      c.cls   = std::get<1>(kv.first);
      c.chain = std::get<2>(kv.first);
      c.work  = work[kv.first];
      for (const auto &t : kv.second) c.columns = c.columns || t.layout == fixedk::LAYOUT_COLUMNS;
      std::vector<uint32_t> items;
      for (uint32_t t = 0; t < kv.second.size(); ++t) {
         if (kv.second[t].numGroups >= (1u << fixedk::kClassGroupBits) || t >= 4096u) return hipErrorInvalidValue;
         for (uint32_t grp = 0; grp < kv.second[t].numGroups; ++grp) items.push_back((t << fixedk::kClassGroupBits) | grp);
      }
      c.numItems = (uint32_t) items.size();
      hipError_t e;
      if ((e = hipMalloc((void **) &c.dTasks, kv.second.size() * sizeof(fixedk::ClassTask))) != hipSuccess) return e;
      h->classes.push_back(c);   // owned by the plan from here on
      if ((e = hipMemcpy(h->classes.back().dTasks, kv.second.data(), kv.second.size() * sizeof(fixedk::ClassTask),
                         hipMemcpyHostToDevice)) != hipSuccess ||
          (e = hipMalloc((void **) &h->classes.back().dItems, items.size() * sizeof(uint32_t))) != hipSuccess ||
          (e = hipMemcpy(h->classes.back().dItems, items.data(), items.size() * sizeof(uint32_t),
                         hipMemcpyHostToDevice)) != hipSuccess)
         return e;
      (void) tuned(class_key(h->classes.back()), c.numItems, &h->classes.back().grid);
   }
   std::sort(h->classes.begin(), h->classes.end(),
             [](const FFTLIB_GPU_ragged_s::Class &a, const FFTLIB_GPU_ragged_s::Class &b) {
                return a.phase != b.phase ? a.phase < b.phase : a.work > b.work;
             });
   return hipSuccess;
}

// Tasks, units and counters of the persistent form: one plain task per
// length up to kRaggedClassMaxN; a column and a row task per transform of a
// longer length (count 1: its offset entry and its scratch region).
hipError_t build_persistent(FFTLIB_GPU_ragged_s *h)
{
   std::vector<fixedk::ClassTask>  tasks;
   std::vector<ragged::QueueEntry> q;
   for (const auto &g : h->groups) {
      if (!g.pTwoPass) {
         fixedk::ClassTask k{};
         k.n = g.n, k.numChannels = g.count, k.layout = fixedk::LAYOUT_PLAIN, k.tw = (const float2 *) g.dTw;
         k.inOff = g.dOff, k.outOff = g.dOff;
         k.numGroups = (k.numChannels + class_group_size(k.n) - 1u) / class_group_size(k.n);
         q.push_back({false, (uint32_t) tasks.size(), k.numGroups, 0u, 0u, g.work});
         tasks.push_back(k);
         continue;
      }
      const uint64_t region = (uint64_t) g.pPlan.n1 * g.pPlan.s2;   // complex points per transform
      const double   work   = g.work / (double) g.count;
      for (uint32_t i = 0; i < g.count; ++i) {
         const fixedk::ClassTask col =
            two_pass_task(g.pPlan, fixedk::LAYOUT_COLUMNS, 1u, g.dOff + i, g.pScrOff + (uint64_t) i * region);
         const fixedk::ClassTask row =
            two_pass_task(g.pPlan, fixedk::LAYOUT_ROWS, 1u, g.dOff + i, g.pScrOff + (uint64_t) i * region);
         q.push_back({true, (uint32_t) tasks.size(), col.numGroups, (uint32_t) tasks.size() + 1u, row.numGroups, work});
         tasks.push_back(col);
         tasks.push_back(row);
      }
   }
   for (const auto &k : tasks)
      if (k.n > kRaggedClassMaxN) return hipErrorInvalidValue;   // excluded when the plan was split
   const std::vector<ragged::QueueUnit> qu = ragged::build_units(q);
   if (qu.empty() || qu.size() > 0xFFFFFFFFull - 65536u) return hipErrorInvalidValue;
   std::vector<fixedk::RaggedUnit> units(qu.size());
   for (size_t i = 0; i < qu.size(); ++i)
      units[i] = fixedk::RaggedUnit{qu[i].task, qu[i].g0, qu[i].g1, qu[i].task2, qu[i].h0, qu[i].h1};
   static_assert(ragged::kQueueNoTask == fixedk::kNoTask, "one marker for a unit without a second task");
   hipError_t e;
   if ((e = hipMalloc((void **) &h->pTasks, tasks.size() * sizeof(fixedk::ClassTask))) != hipSuccess ||
       (e = hipMemcpy(h->pTasks, tasks.data(), tasks.size() * sizeof(fixedk::ClassTask), hipMemcpyHostToDevice)) !=
          hipSuccess ||
       (e = hipMalloc((void **) &h->pUnits, units.size() * sizeof(fixedk::RaggedUnit))) != hipSuccess ||
       (e = hipMemcpy(h->pUnits, units.data(), units.size() * sizeof(fixedk::RaggedUnit), hipMemcpyHostToDevice)) !=
          hipSuccess ||
       (e = hipMalloc((void **) &h->pCounters, fixedk::kRqCounters * sizeof(uint32_t))) != hipSuccess ||
       (e = hipMemset(h->pCounters, 0, fixedk::kRqCounters * sizeof(uint32_t))) != hipSuccess ||
       (e = hipDeviceSynchronize()) != hipSuccess)
      return e;
   h->pNumItems = (uint32_t) units.size();
   return hipSuccess;
}

// Tuning-cache keys of the persistent form: its grid (bucket = item count)
// and the form the tuner chose for plans of a size (bucket = points / 1024).
constexpr uint32_t kPersistentGridKey = 200000u;
constexpr uint32_t kFormKey           = 200001u;
uint32_t           form_count(const FFTLIB_GPU_ragged_s *h)
{
   return (uint32_t) std::min<uint64_t>(h->totalPoints >> 10, 1ull << 30);
}

}  // namespace

extern "C" {

FFTLIB_STATUS FFTLIB_GPU_raggedCreate(FFTLIB_GPU_raggedHandle *handle, const uint32_t *lengths, uint32_t count)
{
   if (handle == nullptr || lengths == nullptr) return FFTLIB_ERR_NULL_POINTER;
   *handle = nullptr;
   if (count == 0) return FFTLIB_ERR_INVALID_DIMENSION;
   std::map<uint32_t, std::vector<uint64_t>> byLen;
   uint64_t                                  off = 0;
   for (uint32_t i = 0; i < count; ++i) {
      const uint32_t n = lengths[i];
      if (n < kRaggedMinN || n > kRaggedMaxN || !smooth235(n)) return FFTLIB_ERR_INVALID_DIMENSION;
      byLen[n].push_back(off);
      off += n;
   }
   auto *h        = new FFTLIB_GPU_ragged_s();
   h->totalPoints = off;
   hipError_t e   = hipSuccess;
   size_t     scr = 0;

   for (auto &kv : byLen) {
      FFTLIB_GPU_ragged_s::Group g;
      g.n     = kv.first;
      g.count = (uint32_t) kv.second.size();
      // lengths up to 16384 are class-kernel tasks; longer ones run two
      // passes through their own scratch region
      g.twoPass = fixed_class_of(g.n) < 0;
      if (g.twoPass) {
         if ((e = two_pass_plan(g.n, &g.plan)) != hipSuccess) {
            release(h);
            delete h;
            return fail("ragged two-pass tables", e);
         }
         g.scrOff = scr / sizeof(float2);
         scr += (two_pass_scratch_points(g.n, g.count) * 8u + 255u) & ~(size_t) 255u;
      }
      // the persistent form splits every length above kRaggedClassMaxN (the
      // same split as the graph form's above 16384); its scratch regions
      // follow the graph form's
      g.pTwoPass = g.n > kRaggedClassMaxN;
      if (g.pTwoPass) {
         if (g.twoPass) g.pPlan = g.plan;
         else if (two_pass_plan(g.n, &g.pPlan) != hipSuccess || g.pPlan.n1 > kRaggedClassMaxN ||
                  g.pPlan.m2 > kRaggedClassMaxN)
            h->pAvailable = false;   // the graph form still runs it in one pass
      }
      if ((e = hipMalloc((void **) &g.dOff, kv.second.size() * sizeof(uint64_t))) == hipSuccess)
         e = hipMemcpy(g.dOff, kv.second.data(), kv.second.size() * sizeof(uint64_t), hipMemcpyHostToDevice);
      if (e == hipSuccess && !g.twoPass) {
         std::vector<float> tw(2 * (size_t) g.n);
         FFTLIB_tw_gen_c2c(tw.data(), (int32_t) g.n);
         if ((e = hipMalloc((void **) &g.dTw, tw.size() * sizeof(float))) == hipSuccess)
            e = hipMemcpy(g.dTw, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice);
      }
      if (e != hipSuccess) {
         h->groups.push_back(g);
         release(h);
         delete h;
         return fail("ragged plan allocation", e);
      }
      double lg = 0.0;
      for (uint32_t t = g.n; t > 1; t >>= 1) lg += 1.0;
      g.work = (double) g.n * lg * g.count;
      h->groups.push_back(g);
   }
   std::sort(h->groups.begin(), h->groups.end(),
             [](const FFTLIB_GPU_ragged_s::Group &a, const FFTLIB_GPU_ragged_s::Group &b) { return a.work > b.work; });
   if ((e = build_classes(h)) != hipSuccess) {
      release(h);
      delete h;
      return fail("ragged class tables", e);
   }
   // persistent form: its own scratch regions in the same buffer (a plan
   // runs one form at a time), 256-byte aligned like the graph form's
   size_t pscr = 0;
   for (auto &g : h->groups)
      if (g.pTwoPass && h->pAvailable) {
         g.pScrOff = pscr / sizeof(float2);
         pscr += ((size_t) g.pPlan.n1 * g.pPlan.s2 * g.count * 8u + 255u) & ~(size_t) 255u;
      }
   scr = std::max(scr, pscr);
   if (h->pAvailable && (e = build_persistent(h)) != hipSuccess) {
      release(h);
      delete h;
      return fail("ragged persistent tables", e);
   }
   {
      uint32_t g = 0;
      if (tuned(kPersistentGridKey, h->pNumItems, &g)) h->pGrid = g;
      if (tuned(kFormKey, form_count(h), &g) &&
          (g == FFTLIB_RAGGED_GRAPH || (g == FFTLIB_RAGGED_PERSISTENT && h->pAvailable)))
         h->form = (int) g;
   }
   if (scr) {
      if ((e = hipMalloc((void **) &h->scratch, scr)) != hipSuccess) {
         release(h);
         delete h;
         return fail("ragged scratch allocation", e);
      }
      h->scratchSize = scr;
   }
   if ((e = hipStreamCreateWithFlags(&h->capture, hipStreamNonBlocking)) != hipSuccess) {
      release(h);
      delete h;
      return fail("hipStreamCreate", e);
   }
   const int lanes = kCaptureLanes;
   for (int i = 0; i < lanes; ++i) {
      hipStream_t s;
      if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) {
         release(h);
         delete h;
         return fail("hipStreamCreate", e);
      }
      h->lanes.push_back(s);
   }
   // the events of the largest call's schedule (forward then inverse)
   const size_t nev = ragged::events_needed(ragged::schedule(kinds_of(h), links_of(h), lanes, 2));
   for (size_t i = 0; i < nev; ++i) {
      hipEvent_t ev;
      if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) {
         release(h);
         delete h;
         return fail("hipEventCreate", e);
      }
      h->events.push_back(ev);
   }
   *handle = h;
   return FFTLIB_SUCCESS;
}

FFTLIB_STATUS FFTLIB_GPU_raggedDestroy(FFTLIB_GPU_raggedHandle handle)
{
   if (handle == nullptr) return FFTLIB_ERR_NULL_POINTER;
   (void) hipStreamSynchronize(handle->capture);
   release(handle);
   delete handle;
   return FFTLIB_SUCCESS;
}

uint64_t FFTLIB_GPU_raggedPoints(FFTLIB_GPU_raggedHandle handle) { return handle ? handle->totalPoints : 0u; }

FFTLIB_STATUS FFTLIB_GPU_raggedExec(FFTLIB_GPU_raggedHandle handle, const FFTLIB_F32 *pX, FFTLIB_F32 *pY,
                                    FFTLIB_F32 *pZ, int32_t direction)
{
   if (handle == nullptr || pX == nullptr || pY == nullptr) return FFTLIB_ERR_NULL_POINTER;
   if (direction < 0 || direction > 2) return FFTLIB_ERR_INVALID_VALUE;
   if (direction == 2 && pZ == nullptr) return FFTLIB_ERR_NULL_POINTER;
   if (((uintptr_t) pX | (uintptr_t) pY | (uintptr_t) (direction == 2 ? pZ : pY)) & 7u)
      return FFTLIB_ERR_NOT_ALIGNED_PTRS_STRIDES;
   std::lock_guard<std::mutex> lock(handle->mu);
   hipError_t                  e;
   if (handle->exec == nullptr || handle->gx != pX || handle->gy != pY || handle->gdir != direction ||
       (direction == 2 && handle->gz != pZ)) {
      // kernels that set attributes or query occupancy do it once, before
      // capture; the warm-up runs after the caller's pending work on pX / pY
      if (handle->exec == nullptr) {
         if ((e = order_after_caller(handle)) != hipSuccess) return fail("ragged stream ordering", e);
         const Direction fwd{pX, pY, C2C_FORWARD};
         if ((e = enqueue(handle, &fwd, 1, handle->capture)) != hipSuccess)
            return fail("ragged warm-up launch", e);
         if ((e = hipStreamSynchronize(handle->capture)) != hipSuccess) return fail("ragged warm-up", e);
      }
      if ((e = capture_graph(handle, pX, pY, pZ, direction)) != hipSuccess) return fail("ragged graph capture", e);
   }
   hipStream_t stream = current_stream();
   if ((e = hipGraphLaunch(handle->exec, stream)) != hipSuccess) return fail("hipGraphLaunch", e);
   if (synchronous() && (e = hipStreamSynchronize(stream)) != hipSuccess) return fail("hipStreamSynchronize", e);
   return FFTLIB_SUCCESS;
}

FFTLIB_STATUS FFTLIB_GPU_raggedTune(FFTLIB_GPU_raggedHandle handle, const FFTLIB_F32 *pX, FFTLIB_F32 *pY)
{
   if (handle == nullptr || pX == nullptr || pY == nullptr) return FFTLIB_ERR_NULL_POINTER;
   std::lock_guard<std::mutex> lock(handle->mu);
   int dev = 0, cus = 0;
   hipError_t e;
   if ((e = hipGetDevice(&dev)) != hipSuccess ||
       (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
      return fail("device query", e);
   hipEvent_t t0, t1;
   if ((e = hipEventCreate(&t0)) != hipSuccess) return fail("hipEventCreate", e);
   if ((e = hipEventCreate(&t1)) != hipSuccess) return fail("hipEventCreate", e);
   hipStream_t s = handle->capture;
   if ((e = order_after_caller(handle)) != hipSuccess) return fail("ragged stream ordering", e);
   // one forward direction of a form, eager, `reps` timed after 1 warm (us)
   const int savedForm  = handle->form;
   auto      time_form = [&](int form, uint32_t grid, int reps, float *us) {
      handle->form = form;
      const uint32_t savedGrid = handle->pGrid;
      handle->pGrid            = grid;
      const Direction fwd{pX, pY, C2C_FORWARD};
      hipError_t      r = enqueue(handle, &fwd, 1, s);
      float           ms = 0.0f;
      if (r == hipSuccess) r = hipEventRecord(t0, s);
      for (int k = 0; k < reps && r == hipSuccess; ++k) r = enqueue(handle, &fwd, 1, s);
      if (r == hipSuccess) r = hipEventRecord(t1, s);
      if (r == hipSuccess) r = hipEventSynchronize(t1);
      if (r == hipSuccess) r = hipEventElapsedTime(&ms, t0, t1);
      handle->pGrid = savedGrid;
      handle->form  = savedForm;
      *us           = ms * 1000.0f / (float) reps;
      return r;
   };
   // graph form: the grid of each class launch by one sweep of coordinate
   // descent on the time of a whole direction (the launches of a phase run
   // side by side, and the grid with which a launch runs fastest alone made
   // the ragged 10000 direction 3% slower than the defaults did), from the
   // grids the plan holds, launches in plan order; a candidate replaces the
   // current grid only when the direction is at least 1% faster with it
   float graphUs = 0.0f;
   e             = time_form(FFTLIB_RAGGED_GRAPH, 0u, 5, &graphUs);
   for (auto &c : handle->classes) {
      if (e != hipSuccess) break;
      // candidates: the default (0: up to 8 x the resident workgroups), then
      // multiples of the CU count (the launch clamps them to its item count)
      const uint32_t u      = (uint32_t) cus;
      const uint32_t cand[] = {0u, 32u * u, 16u * u, 8u * u, 4u * u, 2u * u, u, u / 2u};
      const uint32_t cur    = c.grid;
      uint32_t       bestG  = cur;
      float          best   = graphUs;
      for (uint32_t cap : cand) {
         if (cap == cur) continue;
         c.grid   = cap;
         float us = 0.0f;
         if ((e = time_form(FFTLIB_RAGGED_GRAPH, 0u, 5, &us)) != hipSuccess) break;
         if (us < 0.99f * best) best = us, bestG = cap;
      }
      c.grid = bestG;
      if (e != hipSuccess) break;
      graphUs = best;
      std::lock_guard<std::mutex> tl(g_tuneMu);
      g_tune[TuneKey{class_key(c), bucket_of(c.numItems)}] = TuneVal{bestG, best};
   }
   if (e == hipSuccess) e = time_form(FFTLIB_RAGGED_GRAPH, 0u, 5, &graphUs);
   // persistent form: grid candidates (0 = the resident workgroups)
   float persistUs = 1e30f;
   if (e == hipSuccess && handle->pAvailable) {
      const uint32_t res    = ragged_class_resident();
      const uint32_t cand[] = {0u, res * 3u / 4u, res / 2u, res / 4u};
      uint32_t       bestG  = 0;
      for (uint32_t g : cand) {
         float us = 0.0f;
         if ((e = time_form(FFTLIB_RAGGED_PERSISTENT, g, 3, &us)) != hipSuccess) break;
         if (us < persistUs) persistUs = us, bestG = g;
      }
      if (e == hipSuccess) {
         handle->pGrid = bestG;
         std::lock_guard<std::mutex> tl(g_tuneMu);
         g_tune[TuneKey{kPersistentGridKey, bucket_of(handle->pNumItems)}] = TuneVal{bestG, persistUs};
      }
   }
   // the form: the faster one, unless the caller chose it
   if (e == hipSuccess) {
      const int form = graphUs < persistUs ? FFTLIB_RAGGED_GRAPH : FFTLIB_RAGGED_PERSISTENT;
      if (!handle->formSet) handle->form = form;
      std::lock_guard<std::mutex> tl(g_tuneMu);
      g_tune[TuneKey{kFormKey, bucket_of(form_count(handle))}] = TuneVal{(uint32_t) form, std::min(graphUs, persistUs)};
   }
   (void) hipEventDestroy(t0);
   (void) hipEventDestroy(t1);
   if (e != hipSuccess) return fail("ragged tuning", e);
   // the captured graph (if any) used the old parameters
   if (handle->exec) (void) hipGraphExecDestroy(handle->exec), handle->exec = nullptr;
   if (handle->graph) (void) hipGraphDestroy(handle->graph), handle->graph = nullptr;
   return FFTLIB_SUCCESS;
}

FFTLIB_STATUS FFTLIB_GPU_raggedSetForm(FFTLIB_GPU_raggedHandle handle, int32_t form)
{
   if (handle == nullptr) return FFTLIB_ERR_NULL_POINTER;
   if (form != FFTLIB_RAGGED_GRAPH && form != FFTLIB_RAGGED_PERSISTENT) return FFTLIB_ERR_INVALID_VALUE;
   std::lock_guard<std::mutex> lock(handle->mu);
   if (form == FFTLIB_RAGGED_PERSISTENT && !handle->pAvailable) {
      set_last_error_text("ragged plan: a length's two-pass split has a factor above 4096, which the persistent "
                          "kernel does not hold; the graph form runs it");
      return FFTLIB_ERR_NOT_IMPLEMENTED;
   }
   if (handle->form != form) {   // the captured graph (if any) is of the other form
      if (handle->exec) (void) hipGraphExecDestroy(handle->exec), handle->exec = nullptr;
      if (handle->graph) (void) hipGraphDestroy(handle->graph), handle->graph = nullptr;
   }
   handle->form    = form;
   handle->formSet = true;
   return FFTLIB_SUCCESS;
}

int32_t FFTLIB_GPU_raggedGetForm(FFTLIB_GPU_raggedHandle handle) { return handle ? handle->form : -1; }

FFTLIB_STATUS FFTLIB_GPU_tuneSave(const char *path)
{
   if (path == nullptr) return FFTLIB_ERR_NULL_POINTER;
   FILE *f = fopen(path, "w");
   if (f == nullptr) return FFTLIB_ERR_FAILURE;
   std::lock_guard<std::mutex> lock(g_tuneMu);
   fprintf(f, "fftlib-amd-tune 1\n");
   for (const auto &kv : g_tune)
      fprintf(f, "%u %u %u %.3f\n", kv.first.n, kv.first.bucket, kv.second.grid, (double) kv.second.us);
   const bool ok = fclose(f) == 0;
   return ok ? FFTLIB_SUCCESS : FFTLIB_ERR_FAILURE;
}

FFTLIB_STATUS FFTLIB_GPU_tuneLoad(const char *path)
{
   if (path == nullptr) return FFTLIB_ERR_NULL_POINTER;
   FILE *f = fopen(path, "r");
   if (f == nullptr) return FFTLIB_ERR_FAILURE;
   char magic[32] = {0};
   int  version   = 0;
   if (fscanf(f, "%31s %d", magic, &version) != 2 || strcmp(magic, "fftlib-amd-tune") != 0 || version != 1) {
      fclose(f);
      return FFTLIB_ERR_INVALID_VALUE;
   }
   std::map<TuneKey, TuneVal> loaded;
   unsigned n, b, g;
   double   us;
   int      r;
   while ((r = fscanf(f, "%u %u %u %lf", &n, &b, &g, &us)) == 4) loaded[TuneKey{n, b}] = TuneVal{g, (float) us};
   const bool clean = r == EOF;
   fclose(f);
   if (!clean) return FFTLIB_ERR_INVALID_VALUE;
   std::lock_guard<std::mutex> lock(g_tuneMu);
   for (const auto &kv : loaded) g_tune[kv.first] = kv.second;
   return FFTLIB_SUCCESS;
}

}  // extern "C"
