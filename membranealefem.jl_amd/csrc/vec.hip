// Vector and block-sparse kernels shared by the solvers: BSR SpMV of the
// extended tangent, scaled sums, deterministic reductions (fixed partition and
// fixed summation order, so results do not depend on scheduling) and the
// Newton state update.
#include <algorithm>

#include "internal.h"

namespace mafem {


__global__ __launch_bounds__(256) void k_spmv(int64_t nx, int ndf, const int32_t *__restrict__ rowptr,
                                              const int32_t *__restrict__ colidx,
                                              const double *__restrict__ bsr,
                                              const double *__restrict__ x,
                                              double *__restrict__ y) {
  const int64_t R = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (R >= nx) return;
  const int64_t p = R / ndf;
  const int r = (int)(R % ndf), bb = ndf * ndf;
  double s = 0.0;
  for (int32_t k = rowptr[p]; k < rowptr[p + 1]; ++k) {
    const double *blk = bsr + (int64_t)k * bb + r * ndf;
    const double *xv = x + (int64_t)colidx[k] * ndf;
    for (int c = 0; c < ndf; ++c) s = fma(blk[c], xv[c], s);
  }
  y[R] = s;
}

__global__ __launch_bounds__(256) void k_axpby(int64_t n, double a, const double *__restrict__ x,
                                               double b, double *__restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) y[t] = a * x[t] + b * y[t];
}

constexpr int NRED = 256;  // fixed partition: deterministic reduction order

__global__ __launch_bounds__(256) void k_zero(int64_t n, double *__restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) x[t] = 0.0;
}

__global__ __launch_bounds__(256) void k_sumsq_partial(int64_t n, const double *__restrict__ x,
                                                       const uint8_t *__restrict__ mask,
                                                       double *__restrict__ part) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)NRED * 256) {
    const double v = (mask == nullptr || mask[t]) ? x[t] : 0.0;
    acc = fma(v, v, acc);
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_sum_final(const double *__restrict__ part,
                                                   double *__restrict__ out) {
  __shared__ double s[256];
  s[threadIdx.x] = part[threadIdx.x];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

// cps += Δcps; xms += Δt Δcps[:, motion]  (FiniteElement.jl:41-46)
__global__ __launch_bounds__(256) void k_update_state(int64_t numnp, int ndf, int m0, int m1, int m2,
                                                      const uint8_t *__restrict__ active,
                                                      const double *__restrict__ du, double dt,
                                                      double *__restrict__ cps,
                                                      double *__restrict__ xms) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= numnp) return;
  for (int d = 0; d < ndf; ++d) {
    const int64_t k = p * ndf + d;
    if (active[k]) cps[p + numnp * d] += du[k];
  }
  const int mm[3] = {m0, m1, m2};
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (mm[i] > 0) {
      const int64_t k = p * ndf + (mm[i] - 1);
      if (active[k]) xms[p + numnp * i] += dt * du[k];
    }
}

// ------------------------------------------------------------------ host side

static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

void bsr_spmv(Ctx &c, const double *d_x, double *d_y) {
  hipLaunchKernelGGL(k_spmv, dim3(nblk(c.nx)), dim3(256), 0, c.stream, c.nx, c.ndf, c.d_rowptr,
                     c.d_colidx, c.d_bsr, d_x, d_y);
  MAFEM_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void k_scale(int64_t n, double a, const double *__restrict__ x,
                                               double *__restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) y[t] = a * x[t];
}

void vec_zero(Ctx &c, int64_t n, double *y) {
  hipLaunchKernelGGL(k_zero, dim3(nblk(n)), dim3(256), 0, c.stream, n, y);
  MAFEM_HIP(hipGetLastError());
}

void vec_scale(Ctx &c, int64_t n, double a, const double *x, double *y) {
  hipLaunchKernelGGL(k_scale, dim3(nblk(n)), dim3(256), 0, c.stream, n, a, x, y);
  MAFEM_HIP(hipGetLastError());
}

void vec_axpby(Ctx &c, int64_t n, double a, const double *x, double b, double *y) {
  hipLaunchKernelGGL(k_axpby, dim3(nblk(n)), dim3(256), 0, c.stream, n, a, x, b, y);
  MAFEM_HIP(hipGetLastError());
}

double vec_norm2_masked(Ctx &c, const double *x) {
  hipLaunchKernelGGL(k_sumsq_partial, dim3(NRED), dim3(256), 0, c.stream, c.nx, x, c.d_active,
                     c.d_red);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, c.stream, c.d_red, c.d_red + NRED);
  MAFEM_HIP(hipGetLastError());
  MAFEM_HIP(hipMemcpyAsync(c.h_red, c.d_red + NRED, sizeof(double), hipMemcpyDeviceToHost,
                           c.stream));
  MAFEM_HIP(hipStreamSynchronize(c.stream));
  return sqrt(c.h_red[0]);
}

void apply_update(Ctx &c, const double *d_du, double dt) {
  hipLaunchKernelGGL(k_update_state, dim3(nblk(c.numnp)), dim3(256), 0, c.stream, c.numnp, c.ndf,
                     c.mm_order[0], c.mm_order[1], c.mm_order[2], c.d_active, d_du, dt, c.d_cps,
                     c.d_xms);
  MAFEM_HIP(hipGetLastError());
}

// ------------------------------------------------ multi-vector Krylov kernels

// part[k][b] = partial dot of V_k and w over the fixed chunk of block b
__global__ __launch_bounds__(256) void k_mdot_partial(int64_t n, const double *__restrict__ V,
                                                      int64_t ld, const double *__restrict__ w,
                                                      double *__restrict__ part) {
  __shared__ double s[256];
  const int k = blockIdx.y;
  const double *v = V + k * ld;
  double acc = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)NRED * 256)
    acc = fma(v[t], w[t], acc);
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)k * NRED + blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_mdot_final(const double *__restrict__ part,
                                                    double *__restrict__ out) {
  __shared__ double s[256];
  const int k = blockIdx.x;
  s[threadIdx.x] = part[(int64_t)k * NRED + threadIdx.x];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = s[0];
}

// w += alpha * sum_k coef[k] V_k
__global__ __launch_bounds__(256) void k_mcomb(int64_t n, int nvec, const double *__restrict__ V,
                                               int64_t ld, const double *__restrict__ coef,
                                               double alpha, double *__restrict__ w) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  double s = 0.0;
  for (int k = 0; k < nvec; ++k) s = fma(coef[k], V[k * ld + t], s);
  w[t] = fma(alpha, s, w[t]);
}

double vec_norm2(Ctx &c, const double *x, int64_t n) {
  hipLaunchKernelGGL(k_sumsq_partial, dim3(NRED), dim3(256), 0, c.stream, n, x,
                     (const uint8_t *)nullptr, c.d_red);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, c.stream, c.d_red, c.d_red + NRED);
  MAFEM_HIP(hipGetLastError());
  MAFEM_HIP(hipMemcpyAsync(c.h_red, c.d_red + NRED, sizeof(double), hipMemcpyDeviceToHost,
                           c.stream));
  MAFEM_HIP(hipStreamSynchronize(c.stream));
  return sqrt(c.h_red[0]);
}

// h_out (host) = V^T w for the first nvec columns; blocking
void vec_mdot(Ctx &c, int64_t n, int nvec, const double *V, int64_t ld, const double *w,
              double *h_out) {
  hipLaunchKernelGGL(k_mdot_partial, dim3(NRED, nvec), dim3(256), 0, c.stream, n, V, ld, w,
                     c.d_mdot_part);
  hipLaunchKernelGGL(k_mdot_final, dim3(nvec), dim3(256), 0, c.stream, c.d_mdot_part, c.d_h);
  MAFEM_HIP(hipGetLastError());
  MAFEM_HIP(hipMemcpyAsync(h_out, c.d_h, sizeof(double) * nvec, hipMemcpyDeviceToHost, c.stream));
  MAFEM_HIP(hipStreamSynchronize(c.stream));
}

// w += alpha * V coef, coef given on the host
void vec_mcomb(Ctx &c, int64_t n, int nvec, const double *V, int64_t ld, const double *h_coef,
               double alpha, double *w) {
  MAFEM_HIP(hipMemcpyAsync(c.d_coef, h_coef, sizeof(double) * nvec, hipMemcpyHostToDevice,
                           c.stream));
  hipLaunchKernelGGL(k_mcomb, dim3(nblk(n)), dim3(256), 0, c.stream, n, nvec, V, ld, c.d_coef,
                     alpha, w);
  MAFEM_HIP(hipGetLastError());
}

}  // namespace mafem
