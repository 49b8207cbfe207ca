// Work queue of the persistent ragged launch, as plain host code (no HIP
// types) so that its order and coverage can be checked on a machine without a
// GPU.
//
// One launch runs every transform of one direction; workgroups take units
// from the queue in order. A unit is either
//   - one group of a plain task (a length run whole, many transforms per
//     task), or
//   - one long transform: every group of its column task, then (after a
//     workgroup barrier) every group of its row task. Each long transform has
//     its own column and row task, over its own scratch region, so the row
//     pass reads only what the same workgroup's column pass wrote and no unit
//     waits for another.
// Order: the long transforms, largest first (they are the longest units, so
// they start first and the short plain groups fill the end), then the plain
// groups, largest task first.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

namespace fftlib {
namespace ragged {

constexpr uint32_t kQueueNoTask = 0xFFFFFFFFu;

struct QueueUnit {
   uint32_t task, g0, g1;
   uint32_t task2, h0, h1;   // task2 = kQueueNoTask: none
};

// A plain task (groups of whole transforms), or one long transform: the
// indices of its column and row tasks and their group counts.
struct QueueEntry {
   bool     pair;
   uint32_t task, groups;     // plain: the task; pair: the column task
   uint32_t task2, groups2;   // pair: the row task
   double   work;             // ordering key (larger first)
};

inline std::vector<QueueUnit> build_units(const std::vector<QueueEntry> &e)
{
   std::vector<size_t> pairs, plains;
   for (size_t i = 0; i < e.size(); ++i) (e[i].pair ? pairs : plains).push_back(i);
   auto by_work = [&](size_t a, size_t b) { return e[a].work > e[b].work; };
   std::stable_sort(pairs.begin(), pairs.end(), by_work);
   std::stable_sort(plains.begin(), plains.end(), by_work);
   std::vector<QueueUnit> u;
   for (size_t i : pairs) u.push_back({e[i].task, 0u, e[i].groups, e[i].task2, 0u, e[i].groups2});
   for (size_t i : plains)
      for (uint32_t g = 0; g < e[i].groups; ++g) u.push_back({e[i].task, g, g + 1u, kQueueNoTask, 0u, 0u});
   return u;
}

}  // namespace ragged
}  // namespace fftlib
