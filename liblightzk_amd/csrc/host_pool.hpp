// One pool of host worker threads for the library's host-side parallel
// sections (the MSM window combine, a server step's proof terms and
// assembly). Every caller's loop becomes a job on one shared queue; workers
// take indices from the oldest unfinished job and sleep on a condition
// variable when there is none. Concurrent sections (a server's pipelines, the
// G1 and G2 batches of each) therefore share a fixed set of threads instead
// of each starting its own team, and idle workers do not spin.
#pragma once
#include <functional>

namespace hzk {

// Runs fn(0) .. fn(n - 1) on the pool and the calling thread; returns when
// all have run. Called from a pool worker, it runs the loop inline.
void parallel_for(long n, const std::function<void(long)>& fn);

}  // namespace hzk
