// CPU check of the ragged plan's capture schedule (fftlib_ragged_schedule.h),
// the code the library walks when it captures a ragged call into a graph.
// For many plans (class launches in phase 0 and/or 1, 1..9 lanes, one or two
// directions) it checks:
//  - every event index below events_needed is named exactly once, and the
//    schedule of one direction fits the pool sized for two (the library's
//    pool: events_needed of the two-direction schedule);
//  - the step structure: per direction and phase a fork, the launches of that
//    phase (each class once, lanes dealt round robin), a join;
//  - the ordering the graph must enforce, from a happens-before simulation of
//    event record / stream wait: each phase-1 launch after every phase-0
//    launch of its direction, each launch of the second direction after every
//    launch of the first, and launches of one phase not ordered against each
//    other except on a shared lane.
// Prints "ok" and exits 0 on success; prints the failing plan and exits 1.
#include <stdio.h>
#include <stdlib.h>

#include <set>
#include <vector>

#include "fftlib_ragged_schedule.h"

using namespace fftlib::ragged;

namespace {

int failures = 0;

void fail(const char *what, const std::vector<int> &phases, int lanes, int dirs)
{
   ++failures;
   printf("FAIL %s: lanes %d dirs %d phases", what, lanes, dirs);
   for (int p : phases) printf(" %d", p);
   printf("\n");
}

// Launch node id: dir * classes + cls.
void check(const std::vector<int> &phases, int lanes, int dirs)
{
   const std::vector<Step> s     = schedule(phases, lanes, dirs);
   const size_t            pool2 = events_needed(schedule(phases, lanes, 2), lanes);
   const size_t            need  = events_needed(s, lanes);
   const int               nc    = (int) phases.size();
   if (need > pool2) fail("schedule needs more events than the two-direction pool", phases, lanes, dirs);

   // each event named once, none past `need`
   std::vector<int> uses(need, 0);
   for (const Step &st : s) {
      if (st.op == Op::Fork) {
         if (st.ev >= need) fail("fork event out of range", phases, lanes, dirs);
         else uses[st.ev]++;
      }
      else if (st.op == Op::Join) {
         for (int i = 0; i < lanes; ++i) {
            if (st.ev + i >= need) fail("join event out of range", phases, lanes, dirs);
            else uses[st.ev + i]++;
         }
      }
   }
   for (size_t i = 0; i < need; ++i)
      if (uses[i] != 1) fail("event not named exactly once", phases, lanes, dirs);

   // structure: fork, launches of one phase, join; each class once per direction
   std::vector<std::vector<int>> seen(dirs, std::vector<int>(nc, 0));
   int  open = -1, prevPhase = -1, prevDir = 0, k = 0;
   for (const Step &st : s) {
      if (st.dir < 0 || st.dir >= dirs) {
         fail("bad direction", phases, lanes, dirs);
         return;
      }
      if (st.dir != prevDir) prevPhase = -1, prevDir = st.dir;
      if (st.op == Op::Fork) {
         if (open != -1) fail("fork inside a phase", phases, lanes, dirs);
         open = -2, k = 0;
      }
      else if (st.op == Op::Launch) {
         if (open == -1) fail("launch outside a phase", phases, lanes, dirs);
         if (st.cls < 0 || st.cls >= nc || st.lane < 0 || st.lane >= lanes) {
            fail("launch index out of range", phases, lanes, dirs);
            return;
         }
         const int ph = phases[st.cls];
         if (open == -2) {
            if (ph <= prevPhase) fail("phases out of order", phases, lanes, dirs);
            open = ph;
         }
         else if (ph != open) fail("two phases between one fork and join", phases, lanes, dirs);
         if (st.lane != k++ % lanes) fail("lanes not dealt round robin", phases, lanes, dirs);
         seen[st.dir][st.cls]++;
      }
      else {
         if (open < 0) fail("join without launches", phases, lanes, dirs);
         prevPhase = open, open = -1;
      }
   }
   if (open != -1) fail("schedule ends inside a phase", phases, lanes, dirs);
   for (int d = 0; d < dirs; ++d)
      for (int c = 0; c < nc; ++c)
         if (seen[d][c] != 1) fail("class not launched exactly once per direction", phases, lanes, dirs);

   // happens-before: per stream the set of launches that precede its next
   // operation; an event holds the set of its record point
   const int                     origin = lanes;
   std::vector<std::set<int>>    frontier(lanes + 1);
   std::vector<std::set<int>>    event(need);
   std::vector<std::set<int>>    before(dirs * nc);   // launches ordered before launch (dir, cls)
   std::vector<int>              laneOf(dirs * nc, -1);
   for (const Step &st : s) {
      if (st.op == Op::Fork) {
         event[st.ev] = frontier[origin];
         for (int l = 0; l < lanes; ++l) frontier[l].insert(event[st.ev].begin(), event[st.ev].end());
      }
      else if (st.op == Op::Launch) {
         const int id = st.dir * nc + st.cls;
         before[id]   = frontier[st.lane];
         laneOf[id]   = st.lane;
         frontier[st.lane].insert(id);
      }
      else {
         for (int l = 0; l < lanes; ++l) {
            event[st.ev + l] = frontier[l];
            frontier[origin].insert(event[st.ev + l].begin(), event[st.ev + l].end());
         }
      }
   }
   for (int d = 0; d < dirs; ++d) {
      for (int c = 0; c < nc; ++c) {
         const int id = d * nc + c;
         for (int d2 = 0; d2 <= d; ++d2) {
            for (int c2 = 0; c2 < nc; ++c2) {
               const int  id2  = d2 * nc + c2;
               const bool must = d2 < d || phases[c2] < phases[c];
               const bool is   = before[id].count(id2) != 0;
               if (must && !is) fail("missing order between launches", phases, lanes, dirs);
               if (!must && is && id2 != id && !(d2 == d && laneOf[id2] == laneOf[id]))
                  fail("launches of one phase ordered across lanes", phases, lanes, dirs);
            }
         }
      }
   }
}

}  // namespace

int main()
{
   // every phase pattern of up to 7 class launches (phase 0 launches first, as
   // the plan sorts them), plus unsorted ones
   for (int nc = 0; nc <= 7; ++nc) {
      for (int n0 = 0; n0 <= nc; ++n0) {
         std::vector<int> phases;
         for (int i = 0; i < nc; ++i) phases.push_back(i < n0 ? 0 : 1);
         for (int lanes = 1; lanes <= 9; ++lanes)
            for (int dirs = 1; dirs <= 2; ++dirs) check(phases, lanes, dirs);
      }
   }
   srand(7);
   for (int t = 0; t < 300; ++t) {
      std::vector<int> phases;
      const int        nc = rand() % 12;
      for (int i = 0; i < nc; ++i) phases.push_back(rand() % 2);
      check(phases, 1 + rand() % 9, 1 + rand() % 2);
   }
   // an empty schedule names no events
   if (events_needed(schedule({}, 8, 2), 8) != 0) fail("empty plan needs events", {}, 8, 2);
   if (failures) return 1;
   printf("ok\n");
   return 0;
}
