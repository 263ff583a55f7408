// Kernel plans per FFT/grid length for gfx950, written by pscf_amd/autotune.py
// from timings of the contour sweeps on an MI355X (rerun it to regenerate).
//   max_radix  largest Stockham radix of a transform of this length (8 or 4)
//   xpad       LDS padding (complex words) per packed row of an XY plane
//   zsplit     workgroups sharing one ky row of the z-column step kernel
//   alias      1: the packed-row and half-spectrum LDS buffers of an XY plane
//              share memory (half the LDS per workgroup, more workgroups per CU)
//   ajobs      workgroups per plane of k_step_a: 2 (each repeats the inverse
//              transform and starts one sub-step), 1 (starts both), or 0: 2 while
//              one per plane would leave CUs idle and two fit in one round of
//              resident workgroups, else 1
#pragma once

namespace pscf {

struct TunedPlan {
  int n, max_radix, xpad, zsplit, alias, ajobs;
};

constexpr TunedPlan kTunedPlans[] = {
    {32, 4, 1, 3, 0, 0},  // 29.13 us/step (radix 8, pad 1, 1 row, separate, 2 jobs: 31.60)
    {48, 4, 1, 1, 0, 0},  // 39.63 us/step (radix 8, pad 1, 1 row, separate, 2 jobs: 52.01)
    {64, 4, 1, 1, 0, 0},  // 48.92 us/step (radix 8, pad 1, 1 row, separate, 2 jobs: 51.95)
    {96, 4, 1, 1, 0, 0},  // 129.25 us/step (radix 8, pad 1, 1 row, separate, 2 jobs: 166.22)
};

}  // namespace pscf
