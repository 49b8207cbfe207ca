// Capture schedule of a ragged plan, as plain host code (no HIP types) so
// that its event accounting can be checked on a machine without a GPU.
//
// Each direction is, for each phase that has launches: a fork (one event
// recorded on the origin stream and waited on by every lane), the phase's
// class launches dealt round robin over the lanes, and a join (one event per
// lane, recorded on the lane and waited on by the origin). Every fork and
// join step names the pool events it records, each event once per schedule;
// the plan sizes its pool with events_needed of its largest call (forward
// then inverse) and the capture records exactly the events the steps name.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace fftlib {
namespace ragged {

enum class Op : uint8_t { Fork, Launch, Join };

struct Step {
   Op     op;
   int    dir;    // direction of the step (0 = the first one captured)
   int    cls;    // Launch: index of the class launch in the plan
   int    lane;   // Launch: capture stream it is enqueued on
   size_t ev;     // Fork: its event; Join: the first of `lanes` events (lane i records ev + i)
};

// phases[i]: phase (0 or 1) of class launch i; directions: 1 or 2.
inline std::vector<Step> schedule(const std::vector<int> &phases, int lanes, int directions)
{
   std::vector<Step> s;
   if (lanes < 1) return s;
   size_t ev = 0;
   for (int d = 0; d < directions; ++d) {
      for (int phase = 0; phase < 2; ++phase) {
         bool any = false;
         for (int p : phases) any = any || p == phase;
         if (!any) continue;
         s.push_back({Op::Fork, d, -1, -1, ev});
         ev += 1;
         int next = 0;
         for (int i = 0; i < (int) phases.size(); ++i)
            if (phases[i] == phase) s.push_back({Op::Launch, d, i, next++ % lanes, 0});
         s.push_back({Op::Join, d, -1, -1, ev});
         ev += (size_t) lanes;
      }
   }
   return s;
}

// Size of the event pool a schedule needs: one past the last event it names.
inline size_t events_needed(const std::vector<Step> &s, int lanes)
{
   size_t n = 0;
   for (const Step &st : s) {
      const size_t end = st.op == Op::Fork ? st.ev + 1u : st.op == Op::Join ? st.ev + (size_t) lanes : 0u;
      n                = end > n ? end : n;
   }
   return n;
}

}  // namespace ragged
}  // namespace fftlib
