// Device clock for chare programs: the subset of <cuda/std/chrono> the
// reference examples use (system_clock::now(), time_point difference,
// duration<double>::count() in seconds), on the gfx950 constant-rate wall
// clock.  Replaces the `#include <cuda/std/chrono>` of
// examples/{pingpong,refnum,jacobi2d}/*.h.
#ifndef CHARMING_CHRONO_H_
#define CHARMING_CHRONO_H_

#include <hip/hip_runtime.h>

namespace charm {
__device__ double wall_time();

namespace chrono {

template <class Rep>
struct duration {
  Rep v;
  __device__ duration() : v(0) {}
  __device__ duration(Rep x) : v(x) {}
  __device__ Rep count() const { return v; }
};

template <class Clock>
struct time_point {
  double s;  // seconds since an arbitrary epoch
  __device__ time_point() : s(0) {}
  __device__ explicit time_point(double x) : s(x) {}
  __device__ duration<double> operator-(const time_point& o) const {
    return duration<double>(s - o.s);
  }
};

struct system_clock {
  static __device__ time_point<system_clock> now() {
    return time_point<system_clock>(charm::wall_time());
  }
};

}  // namespace chrono
}  // namespace charm

#endif  // CHARMING_CHRONO_H_
