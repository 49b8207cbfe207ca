// Capture schedule of a ragged plan, as plain host code (no HIP types) so
// that its event accounting and launch ordering can be checked on a machine
// without a GPU.
//
// A schedule is a list of primitive steps on the origin stream and the
// capture lanes: record an event, make a stream wait for an event, launch a
// class kernel. Every event of the pool is recorded at most once per
// schedule; the plan sizes its pool with events_needed of its largest call
// (forward then inverse) and the capture performs exactly these steps.
//
// Class launches are of three kinds: plain (lengths up to 16384 only),
// column (phase 0 launches that hold the column tasks of the longer lengths,
// possibly with plain tasks too) and row. Row tasks are grouped by the
// column launch that holds the column tasks of their transforms (`links`),
// so a row launch reads only what its own column launch wrote. Each
// direction is one of two forms, and every wait in either is at a fork (the
// origin records, every lane waits) or a join (each lane records, the origin
// waits):
//
//   chained (at least one column launch, fewer column launches than lanes,
//   every row launch linked to a column launch): a fork; column launch j on
//   lane j, followed on that lane by the row launches linked to it (stream
//   order is their dependency); the plain launches round robin on the other
//   lanes; a join. The row passes run beside the plain launches' tails.
//
//   phased (otherwise): a fork, the plain and column launches round robin,
//   a join; a fork, the row launches round robin, a join.
//
// The next direction starts after the last join, so it sees every write of
// the previous one (the inverse reads the forward's output and reuses the
// scratch region). Graphs in which lanes waited on other lanes' events
// mid-capture, while further lanes still held unjoined work, crashed the
// host process inside the graph capture or launch; no schedule here has such
// a wait.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace fftlib {
namespace ragged {

enum class Kind : uint8_t { Plain, Column, Row };
enum class Op : uint8_t { Record, Wait, Launch };

// Stream of a step: a lane index, or kOrigin for the stream the call is
// captured on.
constexpr int kOrigin = -1;

struct Step {
   Op     op;
   int    dir;      // direction of the step (0 = the first one captured)
   int    stream;   // lane or kOrigin
   int    cls;      // Launch: index of the class launch in the plan
   size_t ev;       // Record / Wait: pool event
};

namespace detail {

struct Builder {
   std::vector<Step> s;
   size_t            ev = 0;
   int               dir = 0;
   size_t record(int stream)
   {
      s.push_back({Op::Record, dir, stream, -1, ev});
      return ev++;
   }
   void wait(int stream, size_t e) { s.push_back({Op::Wait, dir, stream, -1, e}); }
   void launch(int stream, int cls) { s.push_back({Op::Launch, dir, stream, cls, 0}); }
   void fork(int lanes)
   {
      const size_t e = record(kOrigin);
      for (int l = 0; l < lanes; ++l) wait(l, e);
   }
   void join(int lanes)
   {
      for (int l = 0; l < lanes; ++l) wait(kOrigin, record(l));
   }
};

}  // namespace detail

// kinds[i]: kind of class launch i; links[i]: for a row launch, the index
// of its column launch (else ignored); directions: 1 or 2.
inline std::vector<Step> schedule(const std::vector<Kind> &kinds, const std::vector<int> &links, int lanes,
                                  int directions)
{
   detail::Builder b;
   if (lanes < 1) return b.s;
   const int        n = (int) kinds.size();
   std::vector<int> columnLane(n, -1), plain;
   int              nc = 0;
   for (int i = 0; i < n; ++i) {
      if (kinds[i] == Kind::Column) columnLane[i] = nc++;
      if (kinds[i] == Kind::Plain) plain.push_back(i);
   }
   bool chained = nc >= 1 && nc < lanes && (int) links.size() == n;
   for (int i = 0; chained && i < n; ++i)
      if (kinds[i] == Kind::Row) chained = links[i] >= 0 && links[i] < n && kinds[links[i]] == Kind::Column;
   for (int d = 0; d < directions; ++d) {
      b.dir = d;
      if (n == 0) continue;
      if (chained) {
         b.fork(lanes);
         for (int i = 0; i < n; ++i) {
            if (kinds[i] != Kind::Column) continue;
            b.launch(columnLane[i], i);
            for (int r = 0; r < n; ++r)
               if (kinds[r] == Kind::Row && links[r] == i) b.launch(columnLane[i], r);
         }
         for (size_t k = 0; k < plain.size(); ++k) b.launch(nc + (int) (k % (size_t) (lanes - nc)), plain[k]);
         b.join(lanes);
         continue;
      }
      for (int phase = 0; phase < 2; ++phase) {
         std::vector<int> ls;
         for (int i = 0; i < n; ++i)
            if ((kinds[i] == Kind::Row) == (phase == 1)) ls.push_back(i);
         if (ls.empty()) continue;
         b.fork(lanes);
         for (size_t k = 0; k < ls.size(); ++k) b.launch((int) (k % (size_t) lanes), ls[k]);
         b.join(lanes);
      }
   }
   return b.s;
}

// Size of the event pool a schedule needs: one past the last event it names.
inline size_t events_needed(const std::vector<Step> &s)
{
   size_t n = 0;
   for (const Step &st : s)
      if (st.op != Op::Launch && st.ev + 1u > n) n = st.ev + 1u;
   return n;
}

}  // namespace ragged
}  // namespace fftlib
