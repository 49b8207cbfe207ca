// CPU check of the persistent ragged launch's work queue
// (fftlib_ragged_queue.h): the units workgroups take in order. For many
// random plans (1..60 entries: plain tasks of 1..300 groups and long
// transforms with column and row tasks of 1..300 groups each) it checks:
//  - every group of every plain task is one unit of one group, exactly once;
//  - every long transform is exactly one unit: all groups of its column task,
//    then all groups of its row task (the row pass runs in the workgroup that
//    wrote its input), and no other unit names either task;
//  - the long transforms come first, largest work first, then the plain
//    groups, largest task first, each task's groups in order.
// Prints "ok" and exits 0 on success; prints the failing plans and exits 1.
#include <stdio.h>
#include <stdlib.h>

#include <map>
#include <set>
#include <vector>

#include "fftlib_ragged_queue.h"

using namespace fftlib::ragged;

namespace {

int      failures = 0;
uint64_t rng      = 0x9e3779b97f4a7c15ull;
uint32_t rnd(uint32_t n)
{
   rng ^= rng << 13, rng ^= rng >> 7, rng ^= rng << 17;
   return (uint32_t) (rng % n);
}

void fail(const char *what, int plan)
{
   ++failures;
   if (failures < 20) printf("FAIL plan %d: %s\n", plan, what);
}

void check_plan(int plan)
{
   std::vector<QueueEntry> e;
   uint32_t                next = 0;
   const int               ne   = 1 + (int) rnd(60);
   for (int i = 0; i < ne; ++i) {
      const double w = (double) rnd(20);   // ties on purpose
      if (rnd(2)) e.push_back({false, next++, 1 + rnd(300), 0u, 0u, w});
      else {
         e.push_back({true, next, 1 + rnd(300), next + 1, 1 + rnd(300), w});
         next += 2;
      }
   }
   const std::vector<QueueUnit>           u = build_units(e);
   std::map<uint32_t, const QueueEntry *> byTask;
   for (const auto &x : e) byTask[x.task] = &x;
   std::set<uint64_t> plainSeen;
   std::set<uint32_t> pairSeen, rowTasks;
   size_t             plainGroups = 0, pairs = 0;
   for (const auto &x : e) {
      if (x.pair) rowTasks.insert(x.task2), ++pairs;
      else plainGroups += x.groups;
   }
   bool     inPlain  = false;
   double   lastW    = 1e300;
   uint32_t lastTask = kQueueNoTask, lastG = 0;
   for (const auto &v : u) {
      if (rowTasks.count(v.task)) return fail("row task as a first task", plan);
      auto it = byTask.find(v.task);
      if (it == byTask.end()) return fail("unknown task", plan);
      const QueueEntry &x = *it->second;
      if (x.pair) {
         if (inPlain) return fail("long transform after plain groups", plan);
         if (x.work > lastW) return fail("long transforms not by work", plan);
         lastW = x.work;
         if (v.g0 != 0 || v.g1 != x.groups || v.task2 != x.task2 || v.h0 != 0 || v.h1 != x.groups2)
            return fail("long transform unit not whole", plan);
         if (!pairSeen.insert(v.task).second) return fail("long transform twice", plan);
      }
      else {
         if (!inPlain) inPlain = true, lastW = 1e300;
         if (v.task2 != kQueueNoTask || v.g1 != v.g0 + 1 || v.g0 >= x.groups) return fail("plain unit shape", plan);
         if (v.task != lastTask) {
            if (x.work > lastW) return fail("plain tasks not by work", plan);
            lastW = x.work, lastTask = v.task;
            if (v.g0 != 0) return fail("plain groups not in order", plan);
         }
         else if (v.g0 != lastG + 1) return fail("plain groups not in order", plan);
         lastG = v.g0;
         if (!plainSeen.insert(((uint64_t) v.task << 32) | v.g0).second) return fail("plain group twice", plan);
      }
   }
   if (plainSeen.size() != plainGroups || pairSeen.size() != pairs) return fail("missing units", plan);
}

}  // namespace

int main()
{
   for (int p = 0; p < 2000; ++p) check_plan(p);
   if (!build_units({}).empty()) fail("units from an empty plan", -1);
   if (failures) {
      printf("%d failures\n", failures);
      return 1;
   }
   printf("ok\n");
   return 0;
}
