// tuning.hpp -- launch-shape and threshold choices measured by the
// autotuner (rmetis_amd/autotune.py) and read from tuning.json next to the
// library on first use; every choice is exact (results do not depend on it).
#pragma once

#include <cstdint>

namespace rk {

// A level's bucket: its size, n in [0, 2^19), [2^19, 2^21), [2^21, 2^23),
// [2^23, inf), and its mean degree, m / n in [0, 6), [6, 10), [10, 16),
// [16, 24), [24, 48), [48, inf).
constexpr int kSizeBuckets = 4;
constexpr int kDegBuckets = 6;
constexpr int kLevelBuckets = kSizeBuckets * kDegBuckets;
// Suitor kernel shapes (coarsen.hip): lanes per chain x neighbours per lane
// and round: 0 = 1 x 4, 1 = 2 x 4, 2 = 4 x 2, 3 = 8 x 1.
constexpr int kSuitorShapes = 4;

// The flat table (also the C ABI's layout, rmetis_hip_get_tuning):
//   [0, kLevelBuckets) suitor shape per level bucket, size-major
//                      (size bucket * kDegBuckets + degree bucket)
//   kTuneIncMinN       single-graph levels of at least this many vertices
//                      run incremental cut passes
//   kTuneSpecMin       admission speculates on levels with at least this
//                      many vertices per part
constexpr int kTuneIncMinN = kLevelBuckets;
constexpr int kTuneSpecMin = kLevelBuckets + 1;
constexpr int kTuneCount = kLevelBuckets + 2;

int32_t tune(int index);
// Validates and installs count values (a prefix of the table); false when a
// value is out of range (nothing installed).
bool set_tuning(const int32_t* values, int count, const char** why);
int level_bucket(int64_t n, int64_t m);
// Where the table came from: "built-in defaults" or the file's path.
const char* tuning_source();

}  // namespace rk
