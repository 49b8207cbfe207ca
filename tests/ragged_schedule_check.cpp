// CPU check of the ragged plan's capture schedule (fftlib_ragged_schedule.h),
// the steps the library performs when it captures a ragged call into a
// graph. For many plans (plain, column and row launches in any mix, each
// row launch linked to a column launch or not, 1..9 lanes, one or two
// directions) it checks:
//  - events: each recorded at most once, every wait on an event recorded
//    earlier in the schedule, all below events_needed, and the schedule fits
//    the pool sized for two directions (the library's pool);
//  - launches: each class once per direction, on a lane;
//  - capture shape: every lane's work joined back into the origin stream
//    (each lane's last step is a record that the origin later waits on);
//  - waits only at forks and joins: a lane waits only for events the origin
//    recorded, the origin only for events a lane recorded (graphs with
//    lane-to-lane waits while other lanes held unjoined work crashed the
//    host inside capture or launch);
//  - ordering, from a happens-before simulation of record / wait and stream
//    order: each row launch after its linked column launch (after every
//    column launch when unlinked), every launch of the second direction
//    after every launch of the first; in the chained form (some column
//    launches, fewer than lanes, every row launch linked) no row launch
//    after a plain launch of its direction, and in the phased form every
//    row launch after every launch of phase 0.
// Prints "ok" and exits 0 on success; prints the failing plans and exits 1.
#include <stdio.h>
#include <stdlib.h>

#include <set>
#include <vector>

#include "fftlib_ragged_schedule.h"

using namespace fftlib::ragged;

namespace {

int failures = 0;

void fail(const char *what, const std::vector<Kind> &kinds, int lanes, int dirs)
{
   ++failures;
   printf("FAIL %s: lanes %d dirs %d kinds", what, lanes, dirs);
   for (Kind k : kinds) printf(" %c", k == Kind::Plain ? 'p' : k == Kind::Column ? 'c' : 'r');
   printf("\n");
}

void check(const std::vector<Kind> &kinds, const std::vector<int> &links, int lanes, int dirs)
{
   const std::vector<Step> s     = schedule(kinds, links, lanes, dirs);
   const size_t            need  = events_needed(s);
   const size_t            pool2 = events_needed(schedule(kinds, links, lanes, 2));
   const int               nc    = (int) kinds.size();
   if (need > pool2) fail("schedule needs more events than the two-direction pool", kinds, lanes, dirs);

   // events recorded at most once, waits only on earlier records
   std::vector<int> recorded(need + 1, 0);
   for (const Step &st : s) {
      if (st.op == Op::Launch) {
         if (st.stream < 0 || st.stream >= lanes || st.cls < 0 || st.cls >= nc || st.dir < 0 || st.dir >= dirs) {
            fail("launch index out of range", kinds, lanes, dirs);
            return;
         }
         continue;
      }
      if (st.stream != kOrigin && (st.stream < 0 || st.stream >= lanes)) {
         fail("stream out of range", kinds, lanes, dirs);
         return;
      }
      if (st.ev >= need) {
         fail("event past events_needed", kinds, lanes, dirs);
         return;
      }
      if (st.op == Op::Record && recorded[st.ev]++) fail("event recorded twice", kinds, lanes, dirs);
      if (st.op == Op::Wait && !recorded[st.ev]) fail("wait on an event not yet recorded", kinds, lanes, dirs);
   }
   std::vector<int> recordedOn(need, -2);
   for (const Step &st : s)
      if (st.op == Op::Record) recordedOn[st.ev] = st.stream;
   for (const Step &st : s) {
      if (st.op != Op::Wait) continue;
      const bool laneWaitsOrigin = st.stream != kOrigin && recordedOn[st.ev] == kOrigin;
      const bool originWaitsLane = st.stream == kOrigin && recordedOn[st.ev] != kOrigin;
      if (!laneWaitsOrigin && !originWaitsLane) fail("wait outside a fork or a join", kinds, lanes, dirs);
   }

   // each class launched once per direction
   std::vector<std::vector<int>> seen(dirs, std::vector<int>(nc, 0));
   for (const Step &st : s)
      if (st.op == Op::Launch) seen[st.dir][st.cls]++;
   for (int d = 0; d < dirs; ++d)
      for (int c = 0; c < nc; ++c)
         if (seen[d][c] != 1) fail("class not launched exactly once per direction", kinds, lanes, dirs);

   // every lane with steps ends with a record the origin waits on later
   for (int l = 0; l < lanes; ++l) {
      int last = -1;
      for (int i = 0; i < (int) s.size(); ++i)
         if (s[i].stream == l) last = i;
      if (last < 0) continue;
      bool joined = s[last].op == Op::Record;
      if (joined) {
         joined = false;
         for (int i = last + 1; i < (int) s.size(); ++i)
            if (s[i].op == Op::Wait && s[i].stream == kOrigin && s[i].ev == s[last].ev) joined = true;
      }
      if (!joined) fail("lane work not joined into the origin", kinds, lanes, dirs);
   }

   // happens-before: per stream the launches that precede its next step; an
   // event holds the set at its record
   std::vector<std::set<int>> frontier(lanes + 1);   // index lanes = origin
   std::vector<std::set<int>> event(need);
   std::vector<std::set<int>> before(dirs * nc);
   auto idx = [&](int stream) { return stream == kOrigin ? lanes : stream; };
   for (const Step &st : s) {
      std::set<int> &f = frontier[idx(st.stream)];
      if (st.op == Op::Record) event[st.ev] = f;
      else if (st.op == Op::Wait) f.insert(event[st.ev].begin(), event[st.ev].end());
      else {
         const int id = st.dir * nc + st.cls;
         before[id]   = f;
         f.insert(id);
      }
   }
   int  ncol   = 0;
   bool linked = true;
   for (int c = 0; c < nc; ++c) {
      ncol += kinds[c] == Kind::Column;
      if (kinds[c] == Kind::Row) linked = linked && links[c] >= 0 && links[c] < nc && kinds[links[c]] == Kind::Column;
   }
   const bool chained = ncol >= 1 && ncol < lanes && linked;
   for (int d = 0; d < dirs; ++d) {
      for (int c = 0; c < nc; ++c) {
         const int id = d * nc + c;
         for (int d2 = 0; d2 < dirs; ++d2) {
            for (int c2 = 0; c2 < nc; ++c2) {
               const int  id2 = d2 * nc + c2;
               const bool is  = before[id].count(id2) != 0;
               if (d2 < d && !is) fail("launch not after the previous direction", kinds, lanes, dirs);
               if (d2 > d && is) fail("launch after the next direction", kinds, lanes, dirs);
               if (d2 != d || kinds[c] != Kind::Row) continue;
               const bool myColumn = chained ? links[c] == c2 : kinds[c2] == Kind::Column;
               if (myColumn && !is) fail("row launch not after its column launch", kinds, lanes, dirs);
               if (chained && kinds[c2] == Kind::Plain && is) fail("row launch waits for a plain launch", kinds, lanes, dirs);
               if (!chained && kinds[c2] != Kind::Row && !is) fail("row launch not after phase 0", kinds, lanes, dirs);
            }
         }
      }
   }
}

// Links of the row launches: the first column launch, a random one, or -1
// (no link) when there is no column launch or `unlink`.
std::vector<int> make_links(const std::vector<Kind> &kinds, bool unlink)
{
   std::vector<int> cols, links(kinds.size(), -1);
   for (int i = 0; i < (int) kinds.size(); ++i)
      if (kinds[i] == Kind::Column) cols.push_back(i);
   for (int i = 0; i < (int) kinds.size(); ++i)
      if (kinds[i] == Kind::Row && !cols.empty() && !unlink) links[i] = cols[(size_t) rand() % cols.size()];
   return links;
}

}  // namespace

int main()
{
   // every mix of up to 3 plain, 3 column and 3 row launches in the plan's
   // order (phase 0 first), plus random orders and counts
   for (int np = 0; np <= 3; ++np)
      for (int ncol = 0; ncol <= 3; ++ncol)
         for (int nr = 0; nr <= 3; ++nr)
            for (int lanes = 1; lanes <= 9; ++lanes)
               for (int dirs = 1; dirs <= 2; ++dirs) {
                  std::vector<Kind> k;
                  for (int i = 0; i < ncol; ++i) k.push_back(Kind::Column);
                  for (int i = 0; i < np; ++i) k.push_back(Kind::Plain);
                  for (int i = 0; i < nr; ++i) k.push_back(Kind::Row);
                  check(k, make_links(k, false), lanes, dirs);
                  check(k, make_links(k, true), lanes, dirs);
               }
   srand(7);
   for (int t = 0; t < 600; ++t) {
      std::vector<Kind> k;
      const int         n = rand() % 14;
      for (int i = 0; i < n; ++i) k.push_back((Kind) (rand() % 3));
      check(k, make_links(k, rand() % 5 == 0), 1 + rand() % 9, 1 + rand() % 2);
   }
   if (events_needed(schedule({}, {}, 8, 2)) != 0) fail("empty plan needs events", {}, 8, 2);
   if (failures) return 1;
   printf("ok\n");
   return 0;
}
