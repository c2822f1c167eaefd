// Autotuner of the lanes-per-particle shape of the neighbour-list kernels.
//
// Each tuned kernel runs one particle per group of L adjacent lanes; the best L
// depends on how many neighbours the rows have (a 125-neighbour interior row
// wants wide groups, sparse surface and splash rows narrow ones).  The tuner
// keys on a histogram of the row lengths of the neighbour list the launch reads
// (bucket fractions rounded to 1/16, plus log2 of the row count), times every
// candidate L once per key with HIP events on the library stream, and keeps the
// fastest for the life of the context.  Every tuned kernel computes the same
// result for every L (integer tables in list order; floating-point sums on the
// fixed slot tree of SlotSums), so the choice changes speed only: results stay
// bit-reproducible whatever the timings were.
#pragma once

#include <initializer_list>
#include <vector>

#include "common.hpp"

namespace lsmps {

// Key of a launch of `kernel` over n rows whose neighbour counts are
// offsets[i+1] - offsets[i] (offsets null: the row count alone).  Reads the
// histogram back to the host (one stream synchronisation).
std::vector<int32_t> tune_key(Context& ctx, TunedKernel kernel, size_t n, const int32_t* offsets);

// HIP events around one candidate launch on ctx.stream: tune_mark(ctx, 0)
// before, tune_mark(ctx, 1) after, then tune_elapsed waits and returns ms.
void tune_mark(Context& ctx, int which);
float tune_elapsed(Context& ctx);
// Events and histogram buffer of the tuner (at context creation, so that no
// tuned launch allocates).
void tune_prepare(Context& ctx);

// Runs launch(L) for the lanes chosen among `candidates`: the forced value if
// one is set, else default_lanes without LSMPS_FLAG_AUTOTUNE, else the fastest
// candidate for this launch's key (the first launch of a key runs every
// candidate; they all write the same outputs).  launch enqueues on ctx.stream.
template <typename F>
void launch_tuned(Context& ctx, TunedKernel kernel, std::initializer_list<int> candidates,
                  int default_lanes, size_t n, const int32_t* offsets, F launch) {
  TuneTable& t = ctx.tune;
  int lanes = default_lanes;
  if (t.forced[kernel] > 0) {
    lanes = t.forced[kernel];
  } else if (ctx.autotune() && candidates.size() > 1) {
    const std::vector<int32_t> key = tune_key(ctx, kernel, n, offsets);
    auto it = t.best.find(key);
    if (it != t.best.end()) {
      lanes = it->second;
    } else {
      float best_ms = 0.0f;
      for (int c : candidates) {
        tune_mark(ctx, 0);
        launch(c);
        LSMPS_LAUNCH_CHECK();
        tune_mark(ctx, 1);
        const float ms = tune_elapsed(ctx);
        if (c == *candidates.begin() || ms < best_ms) {
          best_ms = ms;
          lanes = c;
        }
      }
      t.best.emplace(key, lanes);
      ++t.tuned_keys;
      t.last[kernel] = lanes;
      return;  // the candidates' outputs are the result
    }
  }
  bool known = false;
  for (int c : candidates) known = known || c == lanes;
  require(known, "lanes per particle not among the kernel's candidates");
  t.last[kernel] = lanes;
  launch(lanes);
  LSMPS_LAUNCH_CHECK();
}

}  // namespace lsmps
