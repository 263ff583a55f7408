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
    {32, 4, 1, 3},
    {64, 4, 1, 1},
};

}  // namespace pscf
