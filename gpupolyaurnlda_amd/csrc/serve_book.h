// serve_book.h -- host bookkeeping of the fold-in server's published requests (plain C++: the
// server in engine.hip uses it, and a CPU test drives it directly).
//
// A published request is reported by gpulda_serve_poll once its completion flag is raised, or
// as lost once `deadline` seconds have passed since its publication without one (a worker that
// claimed its ring entry and gave up waiting for it, for instance).  The caller then collects it
// with gpulda_serve_result, which returns an error for a lost request; only then does the slot
// enter quarantine.  A quarantined slot keeps its slot number and arena span until no worker
// can run (every pool kernel has ended): a late worker could otherwise still write the span or
// raise the flag of a request that reuses them.  On release the server clears the slot's ring
// entry and record on the device before the slot and span are free again.
#pragma once

#include <cstdint>
#include <vector>

namespace lda {

struct ServeBook {
  std::vector<uint32_t> inflight;    // published, not reported
  std::vector<double> pub_time;      // slot -> host seconds at publication
  std::vector<uint8_t> lost;         // slot -> reported lost, not yet collected or released
  std::vector<uint32_t> quarantine;  // lost and collected: slot and span held until no worker runs
  double deadline = 30.0;            // seconds from publication to a lost report

  void reset(uint32_t n_slots) {
    inflight.clear();
    quarantine.clear();
    pub_time.assign(n_slots, 0.0);
    lost.assign(n_slots, 0);
  }

  void published(uint32_t slot, double now) {
    pub_time[slot] = now;
    lost[slot] = 0;
    inflight.push_back(slot);
  }

  // Moves finished slots (done(slot) true) and, past the deadline, lost ones into out (at most
  // max); returns how many.
  template <class Done>
  uint32_t poll(Done&& done, double now, uint32_t* out, uint32_t max) {
    uint32_t n = 0;
    for (size_t i = 0; i < inflight.size() && n < max;) {
      const uint32_t s = inflight[i];
      bool report = done(s);
      if (!report && now - pub_time[s] > deadline) {
        lost[s] = 1;
        report = true;
      }
      if (report) {
        out[n++] = s;
        inflight[i] = inflight.back();
        inflight.pop_back();
      } else {
        ++i;
      }
    }
    return n;
  }

  // The caller collected a lost slot: it is held until no worker can run.
  void collected_lost(uint32_t slot) {
    lost[slot] = 0;
    quarantine.push_back(slot);
  }

  // Quarantined slots that may be freed now: all of them once no pool kernel runs, else none.
  std::vector<uint32_t> release(bool workers_idle) {
    std::vector<uint32_t> out;
    if (workers_idle) out.swap(quarantine);
    return out;
  }
};

}  // namespace lda
