// gfx950 kernels of the SCFT inner loop (FP64).
//
// Data layout in HBM follows the reference's Fortran arrays (x fastest):
//   real grid      r(x, y, z)            at x + N1*(y + N2*z)             M  = N1*N2*N3
//   half spectrum  k(kx, ky, kz), kx<=N1/2 at kx + N1H*(ky + N2*kz)       Mk = N1H*N2*N3
//   contour slab   q(:, s), s = 0..Ns    at s*M (one slab per chain and direction)
// A 3D r2c transform is one XY-plane pass (x real transform + y transform in
// LDS, one workgroup per z plane) and one Z-column pass (one workgroup per ky
// row of the half spectrum).  The pointwise factors of the contour step are
// fused into those passes, and the z forward transform, the reciprocal-space
// factor and the z inverse transform of one sub-step share a single pass.
#pragma once
#include <hip/hip_runtime.h>

#include "fft_lds.hpp"

namespace pscf {

constexpr int NT = 256;        // threads per workgroup for every kernel below
constexpr int MAXBLK = 16;     // blocks per chain

// One contour sweep (a chain's q or q-dagger) advanced by the step kernels.
// Step j maps contour index in(j) -> out(j): forward in = j, backward
// in = ns - j.  The step interval [lo, lo+1] (lo = min(in, out)) belongs to
// block b with bnd[b] <= lo < bnd[b+1].
struct StepTask {
  double* slab;                 // (ns+1) * M, point s at slab + s*M
  int ns;
  int dir;                      // +1 forward sweep, -1 backward sweep
  int init_one;                 // 1: the first point is the constant 1 (and is written)
  int skip;                     // 1: inactive (converged instance of a batch): every kernel returns
  int nblk;
  int bnd[MAXBLK + 1];
  const double* e1[MAXBLK];     // exp(-ds/2 w)   (M)
  const double* e2[MAXBLK];     // exp(-ds/4 w)   (M)
  const double* k1[MAXBLK];     // exp(-ds b^2 |k|^2 / 6)   (Mk)
  const double* k2[MAXBLK];     // exp(-ds b^2 |k|^2 / 12)  (Mk)
  double2* A;                   // full-step spectrum        (Mk)
  double2* B;                   // first half-step spectrum  (Mk)
  double2* C;                   // second half-step spectrum (Mk)
  double* q1;                   // full-step result          (M)
};

__device__ __forceinline__ int task_in(const StepTask& t, int j) { return t.dir > 0 ? j : t.ns - j; }
__device__ __forceinline__ int task_out(const StepTask& t, int j) { return t.dir > 0 ? j + 1 : t.ns - j - 1; }
__device__ __forceinline__ int task_block(const StepTask& t, int j) {
  const int lo = t.dir > 0 ? j : t.ns - j - 1;
  int b = 0;
  while (b + 1 < t.nblk && lo >= t.bnd[b + 1]) ++b;
  return b;
}

// ----------------------------------------------------------------------------
// Step kernel A (one workgroup per z plane and sweep): finishes step j-1 with
// the second half step's x,y inverse transform, exp(-ds w/4)/M and the
// Richardson combination q = (4 q2 - q1)/3 (stored to the slab), then starts
// step j: exp(-ds w/2) q and exp(-ds w/4) q, each x,y forward transformed.
// Reference: src/step_mod.f90:156-174.
// ----------------------------------------------------------------------------
template <int N1, int N2>
__global__ void __launch_bounds__(NT) k_step_a(const StepTask* __restrict__ tasks, int j, int N3,
                                               const double2* __restrict__ tw1,
                                               const double2* __restrict__ tw2) {
  using PL = Plane<N1, N2>;
  constexpr int PER = (PL::NP * N1 + NT - 1) / NT;
  __shared__ double2 P[PL::PSIZE];
  __shared__ double2 H[PL::HSIZE];
  const int tid = threadIdx.x, z = blockIdx.x;
  const StepTask& t = tasks[blockIdx.y];
  if (t.skip || j > t.ns) return;
  const long M = (long)N1 * N2 * N3;
  const double rM = (double)M;
  const long pbase = (long)z * N1 * N2;
  const long hbase = (long)z * PL::HSIZE;
  double qa[PER], qb[PER];
  if (j == 0) {
    double* qin = t.slab + (long)task_in(t, 0) * M + pbase;
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int i = tid + p * NT;
      if (i < PL::NP * N1) {
        const int x = i % N1, y0 = 2 * (i / N1);
        if (t.init_one) {
          qa[p] = 1.0;
          qb[p] = 1.0;
          qin[y0 * N1 + x] = 1.0;
          if (y0 + 1 < N2) qin[(y0 + 1) * N1 + x] = 1.0;
        } else {
          qa[p] = qin[y0 * N1 + x];
          qb[p] = (y0 + 1 < N2) ? qin[(y0 + 1) * N1 + x] : 0.0;
        }
      }
    }
  } else {
    const int bp = task_block(t, j - 1);
    const double* e2 = t.e2[bp] + pbase;
    const double* q1 = t.q1 + pbase;
    double* qo = t.slab + (long)task_out(t, j - 1) * M + pbase;
    for (int i = tid; i < PL::HSIZE; i += NT) H[i] = t.C[hbase + i];
    __syncthreads();
    plane_inverse<N1, N2, NT>(H, P, tw1, tw2, tid);
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int i = tid + p * NT;
      if (i < PL::NP * N1) {
        const int x = i % N1, y0 = 2 * (i / N1);
        const double2 r = P[i];
        const int o0 = y0 * N1 + x;
        const double q2a = e2[o0] * r.x / rM;
        qa[p] = (4.0 * q2a - q1[o0]) / 3.0;
        qo[o0] = qa[p];
        if (y0 + 1 < N2) {
          const int o1 = o0 + N1;
          const double q2b = e2[o1] * r.y / rM;
          qb[p] = (4.0 * q2b - q1[o1]) / 3.0;
          qo[o1] = qb[p];
        } else {
          qb[p] = 0.0;
        }
      }
    }
    __syncthreads();
  }
  if (j == t.ns) return;
  const int b = task_block(t, j);
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const double* e = (half == 0 ? t.e1[b] : t.e2[b]) + pbase;
    double2* out = (half == 0 ? t.A : t.B) + hbase;
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int i = tid + p * NT;
      if (i < PL::NP * N1) {
        const int x = i % N1, y0 = 2 * (i / N1);
        const int o0 = y0 * N1 + x;
        P[i] = make_double2(e[o0] * qa[p], (y0 + 1 < N2) ? e[o0 + N1] * qb[p] : 0.0);
      }
    }
    __syncthreads();
    plane_forward<N1, N2, NT>(P, H, tw1, tw2, tid);
    for (int i = tid; i < PL::HSIZE; i += NT) out[i] = H[i];
    __syncthreads();
  }
}

// ----------------------------------------------------------------------------
// Step kernel C (two workgroups per z plane and sweep): job 0 turns the full
// step spectrum into q1 = exp(-ds w/2) IFFT[.]/M; job 1 ends the first half
// step and starts the second: exp(-ds w/2) IFFT[.]/M, then the x,y forward
// transform into C.  Reference: src/step_mod.f90:159-167.
// ----------------------------------------------------------------------------
template <int N1, int N2>
__global__ void __launch_bounds__(NT) k_step_c(const StepTask* __restrict__ tasks, int j, int N3,
                                               const double2* __restrict__ tw1,
                                               const double2* __restrict__ tw2) {
  using PL = Plane<N1, N2>;
  __shared__ double2 P[PL::PSIZE];
  __shared__ double2 H[PL::HSIZE];
  const int tid = threadIdx.x, z = blockIdx.x, job = blockIdx.z;
  const StepTask& t = tasks[blockIdx.y];
  if (t.skip || j >= t.ns) return;
  const long M = (long)N1 * N2 * N3;
  const double rM = (double)M;
  const long pbase = (long)z * N1 * N2;
  const long hbase = (long)z * PL::HSIZE;
  const int b = task_block(t, j);
  const double* e1 = t.e1[b] + pbase;
  const double2* src = (job == 0 ? t.A : t.B) + hbase;
  for (int i = tid; i < PL::HSIZE; i += NT) H[i] = src[i];
  __syncthreads();
  plane_inverse<N1, N2, NT>(H, P, tw1, tw2, tid);
  if (job == 0) {
    double* q1 = t.q1 + pbase;
    for (int i = tid; i < PL::NP * N1; i += NT) {
      const int x = i % N1, y0 = 2 * (i / N1);
      const int o0 = y0 * N1 + x;
      const double2 r = P[i];
      q1[o0] = e1[o0] * r.x / rM;
      if (y0 + 1 < N2) q1[o0 + N1] = e1[o0 + N1] * r.y / rM;
    }
    return;
  }
  for (int i = tid; i < PL::NP * N1; i += NT) {
    const int x = i % N1, y0 = 2 * (i / N1);
    const int o0 = y0 * N1 + x;
    const double2 r = P[i];
    P[i] = make_double2(e1[o0] * r.x / rM, (y0 + 1 < N2) ? e1[o0 + N1] * r.y / rM : 0.0);
  }
  __syncthreads();
  plane_forward<N1, N2, NT>(P, H, tw1, tw2, tid);
  double2* out = t.C + hbase;
  for (int i = tid; i < PL::HSIZE; i += NT) out[i] = H[i];
}

// ----------------------------------------------------------------------------
// Z-column convolution (one workgroup per ky row, sweep and job): forward z
// transform, multiply by the reciprocal-space factor, inverse z transform.
// Jobs: kind 0 -> (A, k1) and (B, k2); kind 1 -> (C, k2).
// ----------------------------------------------------------------------------
template <int N1, int N3>
__global__ void __launch_bounds__(NT) k_step_z(const StepTask* __restrict__ tasks, int j, int N2,
                                               int kind, const double2* __restrict__ tw3) {
  constexpr int N1H = N1 / 2 + 1;
  __shared__ double2 T[N1H * N3];
  const int tid = threadIdx.x, ky = blockIdx.x, job = blockIdx.z;
  const StepTask& t = tasks[blockIdx.y];
  if (t.skip || j >= t.ns) return;
  const int b = task_block(t, j);
  double2* arr;
  const double* kf;
  if (kind == 0) {
    arr = job == 0 ? t.A : t.B;
    kf = job == 0 ? t.k1[b] : t.k2[b];
  } else {
    arr = t.C;
    kf = t.k2[b];
  }
  const long rowstride = (long)N1H * N2;
  for (int i = tid; i < N1H * N3; i += NT) {
    const int kz = i / N1H, kx = i % N1H;
    T[i] = arr[kz * rowstride + (long)ky * N1H + kx];
  }
  __syncthreads();
  fft_lines<N3, -1, N1H, 1, N1H, false, NT>(T, tw3, tid);
  for (int i = tid; i < N1H * N3; i += NT) {
    const int kz = i / N1H, kx = i % N1H;
    T[i] = cscale(T[i], kf[kz * rowstride + (long)ky * N1H + kx]);
  }
  __syncthreads();
  fft_lines<N3, +1, N1H, 1, N1H, false, NT>(T, tw3, tid);
  for (int i = tid; i < N1H * N3; i += NT) {
    const int kz = i / N1H, kx = i % N1H;
    arr[kz * rowstride + (long)ky * N1H + kx] = T[i];
  }
}

// ----------------------------------------------------------------------------
// Standalone transform passes (batched over blockIdx.y arrays).
// ----------------------------------------------------------------------------

// real planes -> x,y forward transformed half planes; in[b*istride + ...]
template <int N1, int N2>
__global__ void __launch_bounds__(NT) k_plane_r2c(const double* __restrict__ in, long istride,
                                                  double2* __restrict__ out, long ostride,
                                                  const double2* __restrict__ tw1,
                                                  const double2* __restrict__ tw2) {
  using PL = Plane<N1, N2>;
  __shared__ double2 P[PL::PSIZE];
  __shared__ double2 H[PL::HSIZE];
  const int tid = threadIdx.x, z = blockIdx.x;
  const double* src = in + blockIdx.y * istride + (long)z * N1 * N2;
  for (int i = tid; i < PL::NP * N1; i += NT) {
    const int x = i % N1, y0 = 2 * (i / N1);
    P[i] = make_double2(src[y0 * N1 + x], (y0 + 1 < N2) ? src[(y0 + 1) * N1 + x] : 0.0);
  }
  __syncthreads();
  plane_forward<N1, N2, NT>(P, H, tw1, tw2, tid);
  double2* dst = out + blockIdx.y * ostride + (long)z * PL::HSIZE;
  for (int i = tid; i < PL::HSIZE; i += NT) dst[i] = H[i];
}

// half planes -> x,y inverse transformed real planes, times `scale`
template <int N1, int N2>
__global__ void __launch_bounds__(NT) k_plane_c2r(const double2* __restrict__ in, long istride,
                                                  double* __restrict__ out, long ostride,
                                                  double scale, const double2* __restrict__ tw1,
                                                  const double2* __restrict__ tw2) {
  using PL = Plane<N1, N2>;
  __shared__ double2 P[PL::PSIZE];
  __shared__ double2 H[PL::HSIZE];
  const int tid = threadIdx.x, z = blockIdx.x;
  const double2* src = in + blockIdx.y * istride + (long)z * PL::HSIZE;
  for (int i = tid; i < PL::HSIZE; i += NT) H[i] = src[i];
  __syncthreads();
  plane_inverse<N1, N2, NT>(H, P, tw1, tw2, tid);
  double* dst = out + blockIdx.y * ostride + (long)z * N1 * N2;
  for (int i = tid; i < PL::NP * N1; i += NT) {
    const int x = i % N1, y0 = 2 * (i / N1);
    const double2 r = P[i];
    dst[y0 * N1 + x] = r.x * scale;
    if (y0 + 1 < N2) dst[(y0 + 1) * N1 + x] = r.y * scale;
  }
}

// in-place z transform of W-wide rows (W = N1/2+1 for half spectra, N1 for
// full complex grids), batched over blockIdx.y arrays
template <int W, int N3, int SIGN>
__global__ void __launch_bounds__(NT) k_zfft(double2* __restrict__ data, long stride, int N2,
                                             const double2* __restrict__ tw3) {
  __shared__ double2 T[W * N3];
  const int tid = threadIdx.x, ky = blockIdx.x;
  double2* arr = data + blockIdx.y * stride;
  const long rowstride = (long)W * N2;
  for (int i = tid; i < W * N3; i += NT) T[i] = arr[(i / W) * rowstride + (long)ky * W + i % W];
  __syncthreads();
  fft_lines<N3, SIGN, W, 1, W, false, NT>(T, tw3, tid);
  for (int i = tid; i < W * N3; i += NT) arr[(i / W) * rowstride + (long)ky * W + i % W] = T[i];
}

// full complex planes, in place, x then y (used by the c2c transform)
template <int N1, int N2, int SIGN>
__global__ void __launch_bounds__(NT) k_plane_c2c(double2* __restrict__ data, long stride,
                                                  const double2* __restrict__ tw1,
                                                  const double2* __restrict__ tw2) {
  __shared__ double2 T[N1 * N2];
  const int tid = threadIdx.x, z = blockIdx.x;
  double2* arr = data + blockIdx.y * stride + (long)z * N1 * N2;
  for (int i = tid; i < N1 * N2; i += NT) T[i] = arr[i];
  __syncthreads();
  fft_lines<N1, SIGN, N2, N1, 1, true, NT>(T, tw1, tid);
  fft_lines<N2, SIGN, N1, 1, N1, false, NT>(T, tw2, tid);
  for (int i = tid; i < N1 * N2; i += NT) arr[i] = T[i];
}

}  // namespace pscf
