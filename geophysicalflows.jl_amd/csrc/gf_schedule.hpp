// Radix schedules chosen by the autotuner, per (model, nlayers, nx, ny, dtype), shared by the
// contexts of a process and persisted as a JSON table.
#pragma once
#include <array>
#include <string>

namespace gf {

using SchedKey = std::array<int, 5>;   // model, nlayers, nx, ny, dtype

struct Sched {
  int p[3] = {0, 0, 0};         // FFT points per thread of K1, K2, K3 (0: built-in choice)
  int split[3] = {0, 0, 0};     // K1 one field per workgroup row, K2 u q / v q in two rows,
                                // K3 one column per workgroup
  double step_us = 0;           // measured time of one step under this schedule, microseconds
};

bool sched_lookup(const SchedKey& key, Sched& out);
void sched_store(const SchedKey& key, const Sched& s);
void sched_clear();
// Writes every entry; throws std::runtime_error naming the problem.
void sched_save(const std::string& path);
// Reads a table written by sched_save and stores its entries; returns the number read.
// Throws std::runtime_error on a malformed file or an entry whose schedule is not offered.
int sched_load(const std::string& path);

}  // namespace gf
