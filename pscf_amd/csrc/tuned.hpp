// Kernel plans per FFT/grid length for gfx950, written by pscf_amd/autotune.py
// from timings of the contour sweeps on an MI355X (rerun it to regenerate).
//   max_radix  largest Stockham radix of a transform of this length (8 or 4)
//   xpad       LDS padding (complex words) per packed row of an XY plane
//   zsplit     workgroups sharing one ky row of the z-column step kernel
#pragma once

namespace pscf {

struct TunedPlan {
  int n, max_radix, xpad, zsplit;
};

constexpr TunedPlan kTunedPlans[] = {
    {32, 4, 1, 3},  // 28.63 us/step (radix 8, no pad, 1 row: 31.37)
    {48, 4, 1, 1},  // 39.78 us/step (radix 8, no pad, 1 row: 52.58)
    {64, 4, 1, 1},  // 48.74 us/step (radix 8, no pad, 1 row: 55.27)
    {96, 8, 1, 2},  // 143.22 us/step (radix 8, no pad, 1 row: 157.42)
};

}  // namespace pscf
