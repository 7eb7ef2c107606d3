// Host check of the factor products' static work lists (balanced_lists in
// csrc/layout.cpp), built and run by tests/test_layout_lists.py: every item
// listed exactly once, each worker's items ascending, the worker loads close
// to even, and in spread mode every body's items on distinct workers with
// its items dealt over all 8 XCDs. Exit status 0 when every check holds.
#include <algorithm>
#include <cstdio>
#include <random>
#include <set>
#include <vector>

#include "setup.h"

int main() {
  std::mt19937 rng(7);
  int failures = 0;
  auto fail = [&](const char* what, int trial) {
    std::printf("FAIL %s (trial %d)\n", what, trial);
    ++failures;
  };
  for (int trial = 0; trial < 24; ++trial) {
    const int bodies = 1 + static_cast<int>(rng() % 300);
    const int per = trial % 2 ? 4 : 1;
    const int blocks = 8 * (1 + static_cast<int>(rng() % 100));
    const int workers = blocks * per;
    std::vector<double> cost;
    std::vector<int> body;
    for (int b = 0; b < bodies; ++b) {
      const int n = static_cast<int>(rng() % 700);
      for (int i = 0; i < n; ++i) {
        cost.push_back(1.0 + static_cast<double>(rng() % 9));
        body.push_back(b);
      }
    }
    for (int spread = 0; spread < 2; ++spread) {
      std::vector<int> off, list;
      dphd::balanced_lists(cost, body, bodies, workers, per, off, list, spread != 0);
      if (static_cast<int>(off.size()) != workers + 1 || off[workers] != static_cast<int>(cost.size())) {
        fail("offsets", trial);
        continue;
      }
      std::vector<int> seen(cost.size(), 0), owner(cost.size(), -1);
      for (int w = 0; w < workers; ++w)
        for (int k = off[w]; k < off[w + 1]; ++k) {
          ++seen[list[k]];
          owner[list[k]] = w;
          if (k > off[w] && list[k] <= list[k - 1]) fail("ascending within a worker", trial);
        }
      for (int v : seen)
        if (v != 1) {
          fail("every item exactly once", trial);
          break;
        }
      std::vector<double> load(workers, 0.0);
      double total = 0.0;
      for (size_t i = 0; i < cost.size(); ++i) {
        load[owner[i]] += cost[i];
        total += cost[i];
      }
      const double mean = total / workers, top = *std::max_element(load.begin(), load.end());
      // Graham's bound for the longest-processing-time rule (the busiest worker at most the mean plus
      // one item), with the 5% the XCD split of the bodies may leave between XCDs.
      if (total > 0 && top > 1.05 * mean + 9.0 + 1e-9) fail("load within one item of the mean", trial);
      if (spread) {
        std::vector<std::vector<int>> of(bodies);
        for (size_t i = 0; i < cost.size(); ++i) of[body[i]].push_back(owner[i]);
        for (const auto& v : of) {
          if (static_cast<int>(v.size()) <= workers) {
            const std::set<int> distinct(v.begin(), v.end());
            if (distinct.size() != v.size()) fail("a body's items on distinct workers", trial);
          }
          if (v.size() >= 64) {
            int x[8] = {};
            for (int w : v) ++x[(w / per) % 8];
            for (int k = 0; k < 8; ++k)
              if (x[k] == 0) fail("a body's items on every XCD", trial);
          }
        }
      }
    }
  }
  std::printf("%s\n", failures ? "layout lists: failures" : "layout lists: ok");
  return failures ? 1 : 0;
}
