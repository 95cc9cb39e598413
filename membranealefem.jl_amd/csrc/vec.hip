// Vector and block-sparse kernels shared by the solvers: BSR SpMV of the
// extended tangent, scaled sums, deterministic reductions (fixed partition and
// fixed summation order, so results do not depend on scheduling) and the
// Newton state update.
#include <algorithm>

#include "internal.h"

namespace mafem {


// y = K x over the node blocks (ndf <= 8): eight lanes per row, lane c takes
// column c of every block of the row's node (a wavefront reads whole blocks
// coalesced), then the eight partial sums are added in a fixed tree
__global__ __launch_bounds__(256) void k_spmv(int64_t nx, int ndf, const int32_t *__restrict__ rowptr,
                                              const int32_t *__restrict__ colidx,
                                              const double *__restrict__ bsr,
                                              const double *__restrict__ x,
                                              double *__restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t R = t >> 3;
  const int c = (int)(t & 7);
  if (R >= nx) return;  // whole groups of eight
  const int64_t p = R / ndf;
  const int r = (int)(R % ndf), bb = ndf * ndf;
  double s = 0.0;
  if (c < ndf) {
    for (int32_t k = rowptr[p]; k < rowptr[p + 1]; ++k)
      s = fma(bsr[(int64_t)k * bb + r * ndf + c], x[(int64_t)colidx[k] * ndf + c], s);
  }
  s += __shfl_xor(s, 1, 8);
  s += __shfl_xor(s, 2, 8);
  s += __shfl_xor(s, 4, 8);
  if (c == 0) y[R] = s;
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
  hipLaunchKernelGGL(k_spmv, dim3(nblk(8 * c.nx)), dim3(256), 0, c.stream, c.nx, c.ndf, c.d_rowptr,
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

// ------------------------------------------------------------ batched forms
//
// Vector launches cover nv = nt TILE entries per patch in 256-entry blocks
// (segment B.vec); per-patch norms and dot products are sums of per-block
// partials taken in a fixed order by one workgroup per patch, so they do not
// depend on scheduling.  Every kernel is masked per patch by a flag array.

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double s[256];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  const double r = s[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ int64_t vec_blocks(const PatchDev &d) { return (d.nt * TILE + 255) / 256; }

// sum of the partials of patch p's blocks (one workgroup per patch)
__device__ __forceinline__ double patch_sum(const Seg &sg, const double *__restrict__ part, int p,
                                            int64_t nblk_p) {
  double a = 0.0;
  for (int64_t k = threadIdx.x; k < nblk_p; k += 256) a += part[sg.blk0[p] + k];
  return block_sum(a);
}

struct VecAt {  // patch and entry of a vector-segment thread
  int p;
  int64_t t;
};
__device__ __forceinline__ VecAt vec_at(const Seg &sg) {
  const int p = sg.map[blockIdx.x];
  return {p, (int64_t)(blockIdx.x - sg.blk0[p]) * 256 + threadIdx.x};
}

// b = -r, x = 0; partial |b|^2 over active entries
__global__ __launch_bounds__(256) void kb_solve_begin(Seg sg, const PatchDev *__restrict__ pd,
                                                      const uint8_t *__restrict__ sol,
                                                      double *__restrict__ part) {
  const VecAt at = vec_at(sg);
  double a = 0.0;
  if (sol[at.p]) {
    const PatchDev &d = pd[at.p];
    if (at.t < d.nt * TILE) {
      const double bv = -d.r[at.t];
      d.b[at.t] = bv;
      d.x[at.t] = 0.0;
      if (at.t < d.nx && d.active[at.t]) a = bv * bv;
    }
  }
  a = block_sum(a);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

__global__ __launch_bounds__(256) void kb_solve_begin_fin(Seg sg, const PatchDev *__restrict__ pd,
                                                          BatchFlags f) {
  const int p = blockIdx.x;
  if (!f.on[p]) {
    if (threadIdx.x == 0) f.sol[p] = 0;
    return;
  }
  const double s = patch_sum(sg, f.part, p, vec_blocks(pd[p]));
  if (threadIdx.x == 0) {
    f.bn[p] = sqrt(s);
    f.kits[p] = 0;
    f.relres[p] = 0.0;
    f.sol[p] = s > 0.0 ? 1 : 0;  // a zero right-hand side is solved by x = 0
  }
}

// w = b - K x in FP64 against the block-sparse tangent; partial |w|^2
__global__ __launch_bounds__(256) void kb_resid(Seg sg, const PatchDev *__restrict__ pd,
                                                const uint8_t *__restrict__ mask,
                                                double *__restrict__ part) {
  const VecAt at = vec_at(sg);
  double a = 0.0;
  if (mask[at.p]) {
    const PatchDev &d = pd[at.p];
    const int64_t R = at.t;
    if (R < d.nx) {
      const int64_t q = R / d.ndf;
      const int r = (int)(R % d.ndf), bb = d.ndf * d.ndf;
      double s = 0.0;
      for (int32_t k = d.rowptr[q]; k < d.rowptr[q + 1]; ++k) {
        const double *blk = d.bsr + (int64_t)k * bb + r * d.ndf;
        const double *xv = d.x + (int64_t)d.colidx[k] * d.ndf;
        for (int c = 0; c < d.ndf; ++c) s = fma(blk[c], xv[c], s);
      }
      const double wv = d.b[R] - s;
      d.w[R] = wv;
      if (d.active[R]) a = wv * wv;
    } else if (R < d.nt * TILE) {
      d.w[R] = 0.0;
    }
  }
  a = block_sum(a);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// start of a restart cycle: beta = |w|, g = beta e1, the cycle runs while the
// true relative residual exceeds rtol
__global__ __launch_bounds__(256) void kb_cycle_start(Seg sg, const PatchDev *__restrict__ pd,
                                                      BatchFlags f, int m, double rtol) {
  const int p = blockIdx.x;
  if (!f.sol[p]) return;
  const double s = patch_sum(sg, f.part, p, vec_blocks(pd[p]));
  if (threadIdx.x == 0) {
    const double beta = sqrt(s);
    const double rr = beta / f.bn[p];
    f.relres[p] = rr;
    f.relres0[p] = rr;
    f.kon[p] = rr > rtol ? 1 : 0;
    f.kdim[p] = 0;
    double *g = f.g + (int64_t)p * (m + 1);
    g[0] = beta;
    f.scale[p] = beta > 0.0 ? 1.0 / beta : 0.0;
  }
}

// V_j = scale w
__global__ __launch_bounds__(256) void kb_set_v(Seg sg, const PatchDev *__restrict__ pd,
                                                const uint8_t *__restrict__ mask,
                                                const double *__restrict__ scale, int j) {
  const VecAt at = vec_at(sg);
  if (!mask[at.p]) return;
  const PatchDev &d = pd[at.p];
  const int64_t ld = d.nt * TILE;
  if (at.t < ld) d.V[j * ld + at.t] = scale[at.p] * d.w[at.t];
}

// w = V_j (input of the preconditioner)
__global__ __launch_bounds__(256) void kb_get_v(Seg sg, const PatchDev *__restrict__ pd,
                                                const uint8_t *__restrict__ mask, int j) {
  const VecAt at = vec_at(sg);
  if (!mask[at.p]) return;
  const PatchDev &d = pd[at.p];
  const int64_t ld = d.nt * TILE;
  if (at.t < ld) d.w[at.t] = d.V[j * ld + at.t];
}

// w = K t (padding rows: 0)
__global__ __launch_bounds__(256) void kb_spmv_t(Seg sg, const PatchDev *__restrict__ pd,
                                                 const uint8_t *__restrict__ mask) {
  const VecAt at = vec_at(sg);
  if (!mask[at.p]) return;
  const PatchDev &d = pd[at.p];
  const int64_t R = at.t;
  if (R >= d.nt * TILE) return;
  double s = 0.0;
  if (R < d.nx) {
    const int64_t q = R / d.ndf;
    const int r = (int)(R % d.ndf), bb = d.ndf * d.ndf;
    for (int32_t k = d.rowptr[q]; k < d.rowptr[q + 1]; ++k) {
      const double *blk = d.bsr + (int64_t)k * bb + r * d.ndf;
      const double *tv = d.t + (int64_t)d.colidx[k] * d.ndf;
      for (int c = 0; c < d.ndf; ++c) s = fma(blk[c], tv[c], s);
    }
  }
  d.w[R] = s;
}

// part2[i][block] = partial <V_i, w>, i <= j (blockIdx.y = i)
__global__ __launch_bounds__(256) void kb_mdot(Seg sg, const PatchDev *__restrict__ pd,
                                               const uint8_t *__restrict__ mask,
                                               double *__restrict__ part2) {
  const VecAt at = vec_at(sg);
  const int i = blockIdx.y;
  double a = 0.0;
  if (mask[at.p]) {
    const PatchDev &d = pd[at.p];
    const int64_t ld = d.nt * TILE;
    if (at.t < ld) a = d.V[i * ld + at.t] * d.w[at.t];
  }
  a = block_sum(a);
  if (threadIdx.x == 0) part2[(int64_t)i * sg.nblk + blockIdx.x] = a;
}

// hc[p][i] = <V_i, w>; the Hessenberg column j accumulates it (pass 0 sets)
__global__ __launch_bounds__(256) void kb_mdot_fin(Seg sg, const PatchDev *__restrict__ pd,
                                                   BatchFlags f, int m, int j, int pass) {
  const int p = blockIdx.x, i = blockIdx.y;
  if (!f.kon[p]) return;
  const double s = patch_sum(sg, f.part2 + (int64_t)i * sg.nblk, p, vec_blocks(pd[p]));
  if (threadIdx.x == 0) {
    f.hc[(int64_t)p * (m + 1) + i] = s;
    double &h = f.H[((int64_t)p * m + j) * (m + 1) + i];
    h = pass == 0 ? s : h + s;
  }
}

// w -= sum_{i <= j} hc_i V_i; partial |w|^2
__global__ __launch_bounds__(256) void kb_mcomb_sub(Seg sg, const PatchDev *__restrict__ pd,
                                                    const uint8_t *__restrict__ mask,
                                                    const double *__restrict__ hc, int m, int j,
                                                    double *__restrict__ part) {
  const VecAt at = vec_at(sg);
  double a = 0.0;
  if (mask[at.p]) {
    const PatchDev &d = pd[at.p];
    const int64_t ld = d.nt * TILE;
    if (at.t < ld) {
      const double *h = hc + (int64_t)at.p * (m + 1);
      double s = 0.0;
      for (int i = 0; i <= j; ++i) s = fma(h[i], d.V[i * ld + at.t], s);
      const double wv = d.w[at.t] - s;
      d.w[at.t] = wv;
      a = wv * wv;
    }
  }
  a = block_sum(a);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// end of Arnoldi step j: h_{j+1,j} = |w|, Givens rotations, residual estimate
__global__ __launch_bounds__(256) void kb_arnoldi_end(Seg sg, const PatchDev *__restrict__ pd,
                                                      BatchFlags f, int m, int j, double rtol) {
  const int p = blockIdx.x;
  if (!f.kon[p]) return;
  const double s = patch_sum(sg, f.part, p, vec_blocks(pd[p]));
  if (threadIdx.x == 0) {
    const double hn = sqrt(s);
    double *h = f.H + ((int64_t)p * m + j) * (m + 1);
    double *cs = f.cs + (int64_t)p * m, *sn = f.sn + (int64_t)p * m;
    double *g = f.g + (int64_t)p * (m + 1);
    h[j + 1] = hn;
    for (int i = 0; i < j; ++i) {
      const double a = h[i], b = h[i + 1];
      h[i] = cs[i] * a + sn[i] * b;
      h[i + 1] = -sn[i] * a + cs[i] * b;
    }
    const double a = h[j], b = h[j + 1];
    const double r = hypot(a, b);
    const double c = r > 0.0 ? a / r : 1.0, sv = r > 0.0 ? b / r : 0.0;
    cs[j] = c;
    sn[j] = sv;
    h[j] = r;
    h[j + 1] = 0.0;
    g[j + 1] = -sv * g[j];
    g[j] = c * g[j];
    f.kdim[p] = j + 1;
    f.kits[p] += 1;
    f.kits_tot[p] += 1;
    f.scale[p] = hn > 0.0 ? 1.0 / hn : 0.0;
    // stop at the estimated residual (a little below the target, so the
    // true residual checked after the cycle meets it), a breakdown, or the
    // end of the basis
    if (fabs(g[j + 1]) <= 0.5 * rtol * f.bn[p] || hn == 0.0 || j + 1 == m) f.kon[p] = 0;
  }
}

// least-squares solution y = R^-1 g of the cycle (into hc)
__global__ __launch_bounds__(64) void kb_ysolve(int P, BatchFlags f, int m) {
  const int p = blockIdx.x * 64 + threadIdx.x;
  if (p >= P || !f.sol[p]) return;
  const int k = f.kdim[p];
  const double *g = f.g + (int64_t)p * (m + 1);
  double *y = f.hc + (int64_t)p * (m + 1);
  for (int i = k - 1; i >= 0; --i) {
    double s = g[i];
    for (int l = i + 1; l < k; ++l) s -= f.H[((int64_t)p * m + l) * (m + 1) + i] * y[l];
    y[i] = s / f.H[((int64_t)p * m + i) * (m + 1) + i];
  }
}

// w = sum_{i < kdim} y_i V_i
__global__ __launch_bounds__(256) void kb_vcomb(Seg sg, const PatchDev *__restrict__ pd,
                                                BatchFlags f, int m) {
  const VecAt at = vec_at(sg);
  if (!f.sol[at.p]) return;
  const PatchDev &d = pd[at.p];
  const int64_t ld = d.nt * TILE;
  if (at.t >= ld) return;
  const double *y = f.hc + (int64_t)at.p * (m + 1);
  const int k = f.kdim[at.p];
  double s = 0.0;
  for (int i = 0; i < k; ++i) s = fma(y[i], d.V[i * ld + at.t], s);
  d.w[at.t] = s;
}

// x += t
__global__ __launch_bounds__(256) void kb_add_t(Seg sg, const PatchDev *__restrict__ pd,
                                                const uint8_t *__restrict__ mask) {
  const VecAt at = vec_at(sg);
  if (!mask[at.p]) return;
  const PatchDev &d = pd[at.p];
  if (at.t < d.nt * TILE) d.x[at.t] += d.t[at.t];
}

// after a cycle: the true relative residual decides; converged solves leave sol
__global__ __launch_bounds__(256) void kb_cycle_end(Seg sg, const PatchDev *__restrict__ pd,
                                                    BatchFlags f, double rtol) {
  const int p = blockIdx.x;
  if (!f.sol[p]) return;
  const double s = patch_sum(sg, f.part, p, vec_blocks(pd[p]));
  if (threadIdx.x == 0) {
    const double rr = sqrt(s) / f.bn[p];
    f.relres[p] = rr;
    if (rr <= rtol) f.sol[p] = 0;
  }
}

// count[1] = solves still unconverged
__global__ __launch_bounds__(256) void kb_count_sol(int P, BatchFlags f) {
  int n = 0;
  for (int p = threadIdx.x; p < P; p += 256) n += f.sol[p];
  const double r = block_sum((double)n);
  if (threadIdx.x == 0) f.count[1] = (int32_t)r;
}

// Unconverged solves after the cycles: with a stale factorization, refactor
// and continue from the current iterate; with a fresh FP32 one, move the
// patch to FP64 storage (esc) and continue after refactoring; with a fresh
// FP64 one, refactor with static pivoting (pert, kept for the rest of the
// batch's life) and continue, unless the Newton iteration is diverging (its
// last ε at least kPertMaxEps: a hard failure at once, since no factorization
// helps a garbage tangent); with a fresh FP64 one that already pivots
// statically, continue on the same factors while each cycle at least halves
// the true residual, up to kPertArnoldi Arnoldi steps (its replaced pivots
// leave a few outlying eigenvalues that GMRES needs several restart cycles
// to resolve; a stalled solve, as on the tangent of a diverging Newton
// iteration, stops at once), then keep the inexact correction and count a
// hard failure.
// count[1] = solves to continue, count[2] = factorizations pending,
// count[4] = escalations.
constexpr int32_t kPertArnoldi = 400;
constexpr double kPertMaxEps = 0.1;  // last Newton ε above which the iteration is diverging

__global__ __launch_bounds__(256) void kb_solve_end(int P, const PatchDev *__restrict__ pd,
                                                    BatchFlags f) {
  int redo = 0, pend = 0, hard = 0, esc = 0;
  for (int p = threadIdx.x; p < P; p += 256) {
    if (f.sol[p]) {
      if (!f.fresh[p]) {
        f.fac[p] = 1;
        ++redo;
      } else if (pd[p].prec == 32) {
        f.esc[p] = 1;
        f.fac[p] = 1;
        ++redo;
        ++esc;
      } else if (!f.pert[p] && (f.iters[p] == 0 || f.eps[p] < kPertMaxEps)) {
        // fresh FP64 and the Newton iteration not diverging: refactor once
        // more, with static pivoting
        f.pert[p] = 1;
        f.fac[p] = 1;
        ++redo;
      } else if (f.pert[p] && f.kits[p] < kPertArnoldi && f.relres[p] < 0.5 * f.relres0[p]) {
        // statically pivoted and still converging: another cycle on the same factors
        // (a member that skipped the static-pivoting rung fails at once)
        ++redo;
      } else {
        f.sol[p] = 0;
        ++hard;
      }
    }
    pend += f.on[p] && f.fac[p];
  }
  const double r0 = block_sum((double)redo), r1 = block_sum((double)pend),
               r2 = block_sum((double)hard), r3 = block_sum((double)esc);
  if (threadIdx.x == 0) {
    f.count[1] = (int32_t)r0;
    f.count[2] = (int32_t)r1;
    f.count[3] += (int32_t)r2;
    f.count[4] = (int32_t)r3;
  }
}

__global__ __launch_bounds__(256) void kb_mark_factored(int P, BatchFlags f) {
  for (int p = threadIdx.x; p < P; p += 256)
    if (f.on[p] && f.fac[p]) {
      f.fac[p] = 0;
      f.fresh[p] = 1;
      f.nfac[p] += 1;
    }
}

// refactor before the next solve when this one needed more than
// refactor_iters Arnoldi steps on a factorization that was not fresh
// After a converged solve: a stale factorization that needed more than
// refactor_iters Arnoldi steps is refactored before the next solve; a fresh
// FP32 one that needed more than escalate_iters moves the patch to FP64
// (count[4] = such moves).
__global__ __launch_bounds__(256) void kb_refactor_policy(int P, const PatchDev *__restrict__ pd,
                                                          BatchFlags f, int refactor_iters,
                                                          int escalate_iters) {
  int esc = 0;
  for (int p = threadIdx.x; p < P; p += 256)
    if (f.on[p]) {
      // refactor_iters < 0: a fresh factorization for every Newton iteration
      if (refactor_iters < 0 || (!f.fresh[p] && f.kits[p] > refactor_iters)) f.fac[p] = 1;
      if (f.fresh[p] && pd[p].prec == 32 && f.kits[p] > escalate_iters) {
        f.esc[p] = 1;
        f.fac[p] = 1;
        ++esc;
      }
      f.fresh[p] = 0;
    }
  const double r = block_sum((double)esc);
  if (threadIdx.x == 0) f.count[4] = (int32_t)r;
}

// cps += Δ, xms += Δt Δ[:, motion] (FiniteElement.jl:41-46), one node per thread
__global__ __launch_bounds__(256) void kb_update_state(Seg sg, const PatchDev *__restrict__ pd,
                                                       const uint8_t *__restrict__ on,
                                                       const double *__restrict__ dts) {
  const int p = sg.map[blockIdx.x];
  if (!on[p]) return;
  const PatchDev &d = pd[p];
  const int64_t q = (blockIdx.x - sg.blk0[p]) * 256 + threadIdx.x;
  if (q >= d.numnp) return;
  for (int c = 0; c < d.ndf; ++c) {
    const int64_t k = q * d.ndf + c;
    if (d.active[k]) d.cps[q + d.numnp * c] += d.x[k];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (d.mm[i] > 0) {
      const int64_t k = q * d.ndf + (d.mm[i] - 1);
      if (d.active[k]) d.xms[q + d.numnp * i] += dts[p] * d.x[k];
    }
}

__global__ __launch_bounds__(256) void kb_sumsq_x(Seg sg, const PatchDev *__restrict__ pd,
                                                  const uint8_t *__restrict__ on,
                                                  double *__restrict__ part) {
  const VecAt at = vec_at(sg);
  double a = 0.0;
  if (on[at.p]) {
    const PatchDev &d = pd[at.p];
    if (at.t < d.nx && d.active[at.t]) a = d.x[at.t] * d.x[at.t];
  }
  a = block_sum(a);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// ε = |Δu|_2 / nmdf (FiniteElement.jl:48); a patch stops when ε < εnr
// (:54), when ε is not finite, or after max_iter iterations
__global__ __launch_bounds__(256) void kb_newton_check(Seg sg, const PatchDev *__restrict__ pd,
                                                       BatchFlags f, int max_iter) {
  const int p = blockIdx.x;
  if (!f.on[p]) return;
  const double s = patch_sum(sg, f.part, p, vec_blocks(pd[p]));
  if (threadIdx.x == 0) {
    const double eps = sqrt(s) / pd[p].nmdf;
    const int it = f.iters[p];
    f.eps_hist[(int64_t)p * max_iter + it] = eps;
    f.eps[p] = eps;
    f.iters[p] = it + 1;
    if (!(eps >= pd[p].enr) || it + 1 >= max_iter) f.on[p] = 0;
  }
}

// count[0] = patches still iterating, count[2] = of those, factorizations pending
__global__ __launch_bounds__(256) void kb_count_on(int P, BatchFlags f) {
  int n = 0, m = 0;
  for (int p = threadIdx.x; p < P; p += 256) {
    n += f.on[p];
    m += f.on[p] && f.fac[p];
  }
  const double r = block_sum((double)n), r2 = block_sum((double)m);
  if (threadIdx.x == 0) {
    f.count[0] = (int32_t)r;
    f.count[2] = (int32_t)r2;
  }
}

void b_count(BatchPlan &B) {
  hipLaunchKernelGGL(kb_count_on, dim3(1), dim3(256), 0, B.stream, B.P, B.f);
  MAFEM_HIP(hipGetLastError());
}

void b_solve_begin(BatchPlan &B) {
  const BatchFlags &f = B.f;
  // every patch still iterating solves: sol is first set from on
  hipLaunchKernelGGL(kb_solve_begin, dim3(B.vec.nblk), dim3(256), 0, B.stream, B.vec, B.d_pd, f.on,
                     f.part);
  hipLaunchKernelGGL(kb_solve_begin_fin, dim3(B.P), dim3(256), 0, B.stream, B.vec, B.d_pd, f);
  MAFEM_HIP(hipGetLastError());
}

void b_gmres_cycle(BatchPlan &B) {
  const BatchFlags &f = B.f;
  const unsigned nv = B.vec.nblk;
  hipLaunchKernelGGL(kb_resid, dim3(nv), dim3(256), 0, B.stream, B.vec, B.d_pd, f.sol, f.part);
  hipLaunchKernelGGL(kb_cycle_start, dim3(B.P), dim3(256), 0, B.stream, B.vec, B.d_pd, f, B.m,
                     B.rtol);
  hipLaunchKernelGGL(kb_set_v, dim3(nv), dim3(256), 0, B.stream, B.vec, B.d_pd, f.kon, f.scale, 0);
  for (int j = 0; j < B.m; ++j) {
    hipLaunchKernelGGL(kb_get_v, dim3(nv), dim3(256), 0, B.stream, B.vec, B.d_pd, f.kon, j);
    b_band_solve(B, f.kon);
    hipLaunchKernelGGL(kb_spmv_t, dim3(nv), dim3(256), 0, B.stream, B.vec, B.d_pd, f.kon);
    for (int pass = 0; pass < 2; ++pass) {  // classical Gram-Schmidt, twice
      hipLaunchKernelGGL(kb_mdot, dim3(nv, j + 1), dim3(256), 0, B.stream, B.vec, B.d_pd, f.kon,
                         f.part2);
      hipLaunchKernelGGL(kb_mdot_fin, dim3(B.P, j + 1), dim3(256), 0, B.stream, B.vec, B.d_pd, f,
                         B.m, j, pass);
      hipLaunchKernelGGL(kb_mcomb_sub, dim3(nv), dim3(256), 0, B.stream, B.vec, B.d_pd, f.kon, f.hc,
                         B.m, j, f.part);
    }
    hipLaunchKernelGGL(kb_arnoldi_end, dim3(B.P), dim3(256), 0, B.stream, B.vec, B.d_pd, f, B.m, j,
                       B.rtol);
    if (j + 1 < B.m)
      hipLaunchKernelGGL(kb_set_v, dim3(nv), dim3(256), 0, B.stream, B.vec, B.d_pd, f.kon, f.scale,
                         j + 1);
    MAFEM_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(kb_ysolve, dim3((B.P + 63) / 64), dim3(64), 0, B.stream, B.P, f, B.m);
  hipLaunchKernelGGL(kb_vcomb, dim3(nv), dim3(256), 0, B.stream, B.vec, B.d_pd, f, B.m);
  b_band_solve(B, f.sol);
  hipLaunchKernelGGL(kb_add_t, dim3(nv), dim3(256), 0, B.stream, B.vec, B.d_pd, f.sol);
  hipLaunchKernelGGL(kb_resid, dim3(nv), dim3(256), 0, B.stream, B.vec, B.d_pd, f.sol, f.part);
  hipLaunchKernelGGL(kb_cycle_end, dim3(B.P), dim3(256), 0, B.stream, B.vec, B.d_pd, f, B.rtol);
  hipLaunchKernelGGL(kb_count_sol, dim3(1), dim3(256), 0, B.stream, B.P, f);
  MAFEM_HIP(hipGetLastError());
}

void b_solve_end(BatchPlan &B) {
  hipLaunchKernelGGL(kb_solve_end, dim3(1), dim3(256), 0, B.stream, B.P, B.d_pd, B.f);
  MAFEM_HIP(hipGetLastError());
}

void b_mark_factored(BatchPlan &B) {
  hipLaunchKernelGGL(kb_mark_factored, dim3(1), dim3(256), 0, B.stream, B.P, B.f);
  MAFEM_HIP(hipGetLastError());
}

void b_update(BatchPlan &B) {
  hipLaunchKernelGGL(kb_refactor_policy, dim3(1), dim3(256), 0, B.stream, B.P, B.d_pd, B.f,
                     B.refactor_iters, B.escalate_iters);
  hipLaunchKernelGGL(kb_update_state, dim3(B.node.nblk), dim3(256), 0, B.stream, B.node, B.d_pd,
                     B.f.on, B.f.dt);
  hipLaunchKernelGGL(kb_sumsq_x, dim3(B.vec.nblk), dim3(256), 0, B.stream, B.vec, B.d_pd, B.f.on,
                     B.f.part);
  hipLaunchKernelGGL(kb_newton_check, dim3(B.P), dim3(256), 0, B.stream, B.vec, B.d_pd, B.f,
                     B.max_iter);
  MAFEM_HIP(hipGetLastError());
  b_count(B);
}

}  // namespace mafem
