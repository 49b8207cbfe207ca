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
// Plan: transforms are grouped by length. A group whose length has a
// specialised kernel (every 2/3/5-smooth length 16..16384) runs it over the
// group's channel offset table; longer groups take the outer-pass + LDS-block
// path with a plan-owned scratch buffer. The groups of one direction are
// spread over several capture streams (forked from and joined into the
// capture origin), so the graph lets independent groups run concurrently;
// a forward-then-inverse call joins between the two directions.
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "fft_kernels.h"
#include "fftlib_gpu.h"
#include "fftlib_runtime.h"

using namespace fftlib;

struct FFTLIB_GPU_ragged_s {
   struct Group {
      uint32_t          n        = 0;
      uint32_t          count    = 0;
      bool              large    = false;   // longer than the specialised kernels
      StageTable        st{};
      uint64_t         *dOff     = nullptr; // channel offsets (complex elements)
      float            *dTw      = nullptr; // twiddle table of length n
      uint32_t          maxGrid  = 0;       // tuned persistent grid cap (0 = resident slots)
      double            work     = 0.0;     // n * log2(n) * count, for ordering
   };
   std::vector<Group> groups;
   uint64_t           totalPoints = 0;
   float             *scratch     = nullptr;
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
// Key: length and a power-of-two bucket of the group size.
struct TuneKey {
   uint32_t n, bucket;
   bool     operator<(const TuneKey &o) const { return n != o.n ? n < o.n : bucket < o.bucket; }
};
struct TuneVal {
   uint32_t maxGrid;
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

bool tuned(uint32_t n, uint32_t count, uint32_t *maxGrid)
{
   std::lock_guard<std::mutex> lock(g_tuneMu);
   auto it = g_tune.find(TuneKey{n, bucket_of(count)});
   if (it == g_tune.end()) return false;
   *maxGrid = it->second.maxGrid;
   return true;
}

hipError_t launch_group(const FFTLIB_GPU_ragged_s::Group &g, const float *in, float *out, int mode, float *scratch,
                        uint32_t maxGrid, hipStream_t s)
{
   if (g.large)
      return launch_c2c_large(in, out, g.dTw, g.st, g.count, mode, scratch, s, g.dOff);
   return launch_c2c_fixed_ragged(in, out, g.dTw, g.n, g.count, mode, g.dOff, maxGrid, s);
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
   if (h->scratch) (void) hipFree(h->scratch);
   for (hipEvent_t e : h->events) (void) hipEventDestroy(e);
   for (hipStream_t s : h->lanes) (void) hipStreamDestroy(s);
   if (h->capture) (void) hipStreamDestroy(h->capture);
}

// Enqueues one direction: groups dealt to the lanes largest first.
hipError_t enqueue_direction(FFTLIB_GPU_ragged_s *h, const float *in, float *out, int mode, hipStream_t origin,
                             size_t &ev)
{
   hipError_t e;
   hipEvent_t fork = h->events[ev++];
   if ((e = hipEventRecord(fork, origin)) != hipSuccess) return e;
   for (hipStream_t lane : h->lanes)
      if ((e = hipStreamWaitEvent(lane, fork, 0)) != hipSuccess) return e;
   // The large groups share the plan's scratch, so they run in order on lane
   // 0; the other groups are dealt round-robin to the remaining lanes.
   const size_t nl   = h->lanes.size();
   size_t       next = 0;
   for (const auto &g : h->groups) {
      hipStream_t lane = h->lanes[0];
      if (!g.large && nl > 1) lane = h->lanes[1 + (next++ % (nl - 1))];
      if ((e = launch_group(g, in, out, mode, h->scratch, g.maxGrid, lane)) != hipSuccess) return e;
   }
   for (hipStream_t lane : h->lanes) {
      hipEvent_t join = h->events[ev++];
      if ((e = hipEventRecord(join, lane)) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(origin, join, 0)) != hipSuccess) return e;
   }
   return hipSuccess;
}

hipError_t capture_graph(FFTLIB_GPU_ragged_s *h, const float *x, float *y, float *z, int dir)
{
   if (h->exec) (void) hipGraphExecDestroy(h->exec), h->exec = nullptr;
   if (h->graph) (void) hipGraphDestroy(h->graph), h->graph = nullptr;
   hipError_t e = hipStreamBeginCapture(h->capture, hipStreamCaptureModeThreadLocal);
   if (e != hipSuccess) return e;
   size_t     ev = 0;
   hipError_t le = hipSuccess;
   if (dir == 0) le = enqueue_direction(h, x, y, C2C_FORWARD, h->capture, ev);
   else if (dir == 1) le = enqueue_direction(h, x, y, C2C_INVERSE_PRESCALE, h->capture, ev);
   else {
      le = enqueue_direction(h, x, y, C2C_FORWARD, h->capture, ev);
      if (le == hipSuccess) le = enqueue_direction(h, y, z, C2C_INVERSE_PRESCALE, h->capture, ev);
   }
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
      if (!build_stage_table(g.n, &g.st)) {
         release(h);
         delete h;
         return FFTLIB_ERR_INVALID_DIMENSION;
      }
      // every 2/3/5-smooth length from 16 to 16384 has a specialised kernel;
      // longer ones take the outer-pass path and share the plan's scratch
      g.large = !fixed_c2c_supported(g.n);
      if (g.large) scr = std::max(scr, (size_t) g.n * g.count * 8u);
      std::vector<float> tw(2 * (size_t) g.n);
      FFTLIB_tw_gen_c2c(tw.data(), (int32_t) g.n);
      if ((e = hipMalloc((void **) &g.dOff, kv.second.size() * sizeof(uint64_t))) != hipSuccess ||
          (e = hipMemcpy(g.dOff, kv.second.data(), kv.second.size() * sizeof(uint64_t), hipMemcpyHostToDevice)) !=
             hipSuccess ||
          (e = hipMalloc((void **) &g.dTw, tw.size() * sizeof(float))) != hipSuccess ||
          (e = hipMemcpy(g.dTw, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) {
         h->groups.push_back(g);
         release(h);
         delete h;
         return fail("ragged plan allocation", e);
      }
      double lg = 0.0;
      for (uint32_t t = g.n; t > 1; t >>= 1) lg += 1.0;
      g.work = (double) g.n * lg * g.count;
      (void) tuned(g.n, g.count, &g.maxGrid);
      h->groups.push_back(g);
   }
   std::sort(h->groups.begin(), h->groups.end(),
             [](const FFTLIB_GPU_ragged_s::Group &a, const FFTLIB_GPU_ragged_s::Group &b) { return a.work > b.work; });
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
   const int lanes = (int) std::min<size_t>(kCaptureLanes, h->groups.size());
   for (int i = 0; i < lanes; ++i) {
      hipStream_t s;
      if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) {
         release(h);
         delete h;
         return fail("hipStreamCreate", e);
      }
      h->lanes.push_back(s);
   }
   // fork + joins for up to two directions
   const size_t nev = 2 * (1 + h->lanes.size());
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
         for (const auto &g : handle->groups) {
            if ((e = launch_group(g, pX, pY, C2C_FORWARD, handle->scratch, g.maxGrid, handle->capture)) !=
                hipSuccess)
               return fail("ragged warm-up launch", e);
         }
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
   for (auto &g : handle->groups) {
      if (g.large) continue;   // the large path has no grid parameter
      // candidates: resident slots (0), then caps at multiples of the CU count
      const uint32_t cand[] = {0u, 4u * (uint32_t) cus, 2u * (uint32_t) cus, (uint32_t) cus, (uint32_t) cus / 2u};
      float          best   = 1e30f;
      uint32_t       bestG  = 0;
      for (uint32_t c : cand) {
         if ((e = launch_group(g, pX, pY, C2C_FORWARD, handle->scratch, c, s)) != hipSuccess) break;  // warm
         float ms = 0.0f;
         if ((e = hipEventRecord(t0, s)) != hipSuccess) break;
         for (int r = 0; r < 3 && e == hipSuccess; ++r) e = launch_group(g, pX, pY, C2C_FORWARD, handle->scratch, c, s);
         if (e != hipSuccess) break;
         if ((e = hipEventRecord(t1, s)) != hipSuccess || (e = hipEventSynchronize(t1)) != hipSuccess ||
             (e = hipEventElapsedTime(&ms, t0, t1)) != hipSuccess)
            break;
         if (ms < best) best = ms, bestG = c;
      }
      if (e != hipSuccess) break;
      g.maxGrid = bestG;
      std::lock_guard<std::mutex> tl(g_tuneMu);
      g_tune[TuneKey{g.n, bucket_of(g.count)}] = TuneVal{bestG, best * 1000.0f / 3.0f};
   }
   (void) hipEventDestroy(t0);
   (void) hipEventDestroy(t1);
   if (e != hipSuccess) return fail("ragged tuning", e);
   // the captured graph (if any) used the old parameters
   if (handle->exec) (void) hipGraphExecDestroy(handle->exec), handle->exec = nullptr;
   if (handle->graph) (void) hipGraphDestroy(handle->graph), handle->graph = nullptr;
   return FFTLIB_SUCCESS;
}

FFTLIB_STATUS FFTLIB_GPU_tuneSave(const char *path)
{
   if (path == nullptr) return FFTLIB_ERR_NULL_POINTER;
   FILE *f = fopen(path, "w");
   if (f == nullptr) return FFTLIB_ERR_FAILURE;
   std::lock_guard<std::mutex> lock(g_tuneMu);
   fprintf(f, "fftlib-amd-tune 1\n");
   for (const auto &kv : g_tune)
      fprintf(f, "%u %u %u %.3f\n", kv.first.n, kv.first.bucket, kv.second.maxGrid, (double) kv.second.us);
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
