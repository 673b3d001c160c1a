// Drives the fold-in server's request bookkeeping (gpupolyaurnlda_amd/csrc/serve_book.h) through
// its paths on the CPU; exits nonzero on the first failed check.
#include <algorithm>
#include <cstdio>
#include <set>
#include <vector>

#include "serve_book.h"

#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
      return 1;                                                    \
    }                                                              \
  } while (0)

int main() {
  lda::ServeBook b;
  b.reset(8);
  b.deadline = 10.0;
  std::set<unsigned> done;
  auto is_done = [&](unsigned s) { return done.count(s) != 0; };
  unsigned out[8];
  for (unsigned s : {0u, 1u, 2u, 3u}) b.published(s, 0.0);
  // nothing finished yet, before the deadline: nothing reported
  CHECK(b.poll(is_done, 5.0, out, 8) == 0 && b.inflight.size() == 4);
  // a finished slot is reported once
  done.insert(1);
  CHECK(b.poll(is_done, 6.0, out, 8) == 1 && out[0] == 1 && !b.lost[1]);
  CHECK(b.poll(is_done, 6.5, out, 8) == 0);
  // past the deadline the unfinished ones are reported lost, at most max per call
  CHECK(b.poll(is_done, 10.5, out, 2) == 2);
  std::vector<unsigned> first(out, out + 2);
  CHECK(b.poll(is_done, 10.6, out, 8) == 1);
  first.push_back(out[0]);
  std::sort(first.begin(), first.end());
  CHECK((first == std::vector<unsigned>{0, 2, 3}));
  CHECK(b.lost[0] && b.lost[2] && b.lost[3] && b.inflight.empty());
  // a lost slot is held until it is collected and no worker runs
  CHECK(b.release(true).empty());
  b.collected_lost(2);
  CHECK(!b.lost[2] && b.quarantine.size() == 1);
  CHECK(b.release(false).empty() && b.quarantine.size() == 1);
  std::vector<unsigned> rel = b.release(true);
  CHECK(rel.size() == 1 && rel[0] == 2 && b.quarantine.empty());
  // a released slot is published afresh: reported normally, never as lost from its old life
  b.published(2, 20.0);
  done.insert(2);
  CHECK(b.poll(is_done, 21.0, out, 8) == 1 && out[0] == 2 && !b.lost[2]);
  // a slot finishing exactly when the deadline passes is reported as finished, not lost
  b.published(5, 30.0);
  done.insert(5);
  CHECK(b.poll(is_done, 45.0, out, 8) == 1 && out[0] == 5 && !b.lost[5]);
  std::puts("serve book ok");
  return 0;
}
