// Direct solve of the Newton system on gfx950: banded LU of the extended
// node-block system, stored as dense 64x64 tiles inside the band.
//
// Natural node numbering (i1 + num1np i2) bounds the coupling of two nodes
// that share an element to |Δnode| <= 2 num1np + 2, so the scalar matrix has
// half bandwidth kl = (2 num1np + 3) ndf - 1 and LU without pivoting across
// tile rows keeps the fill inside that band.  Partial pivoting is applied
// inside each diagonal tile.  Per tile step K:
//   k_diag   : P_K A_KK = L_KK U_KK in one wavefront (one row per lane), then
//              explicit inverses of L_KK and U_KK;
//   k_panel  : U_KJ = L_KK^-1 P_K A_KJ, L_IK = A_IK U_KK^-1, and P_K applied to
//              the L tiles already stored in tile row K;
//   k_update : A_IJ -= L_IK U_KJ over the (tb x tb) trailing window.
// This replaces `K_gl \ r_gl` (src/analysis/FiniteElement.jl:38, UMFPACK).
#include "internal.h"

namespace mafem {

constexpr int T = TILE;

__device__ __forceinline__ double *tile_ptr(double *band, int tb, int64_t I, int64_t J) {
  return band + ((I * (2 * tb + 1) + (J - I + tb)) * (int64_t)T * T);
}

// ----------------------------------------------------------------- band fill

__global__ __launch_bounds__(256) void k_zero(int64_t n, double *__restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) x[t] = 0.0;
}

__global__ __launch_bounds__(256) void k_bsr_to_band(
    int64_t nnzb, int ndf, int tb, const int32_t *__restrict__ brow,
    const int32_t *__restrict__ colidx, const double *__restrict__ bsr,
    double *__restrict__ band) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int bb = ndf * ndf;
  if (t >= nnzb * bb) return;
  const int64_t blk = t / bb;
  const int rc = (int)(t % bb), r = rc / ndf, c = rc % ndf;
  const int64_t R = (int64_t)brow[blk] * ndf + r, C = (int64_t)colidx[blk] * ndf + c;
  const int64_t I = R / T, J = C / T;
  tile_ptr(band, tb, I, J)[(R % T) * T + (C % T)] = bsr[t];
}

// identity on the padding rows beyond the extended system
__global__ void k_band_pad(int64_t nx, int64_t npad, int tb, double *__restrict__ band) {
  const int64_t R = nx + (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (R >= npad) return;
  const int64_t I = R / T;
  tile_ptr(band, tb, I, I)[(R % T) * T + (R % T)] = 1.0;
}

// ------------------------------------------------------------ factorization

__device__ __forceinline__ void wave_argmax(double &v, int &idx) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off, 64);
    const int oi = __shfl_xor(idx, off, 64);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
}

// One wavefront factors the 64x64 diagonal tile with partial pivoting.
__global__ __launch_bounds__(64) void k_diag(int64_t K, int tb, double *__restrict__ band,
                                             double *__restrict__ linv,
                                             double *__restrict__ uinv,
                                             int32_t *__restrict__ piv) {
  __shared__ double sP[T];            // broadcast pivot row
  __shared__ double sA[T], sB2[T];    // row exchange
  __shared__ double sT[T][T + 1];     // factored tile for the inverses
  const int lane = threadIdx.x;
  double *A = tile_ptr(band, tb, K, K);
  double r[T];
#pragma unroll
  for (int j = 0; j < T; ++j) r[j] = A[lane * T + j];
  int perm = lane;

#pragma unroll
  for (int j = 0; j < T; ++j) {
    double v = (lane >= j) ? fabs(r[j]) : -1.0;
    int p = lane;
    wave_argmax(v, p);
    if (p != j) {
      if (lane == p) {
#pragma unroll
        for (int k = 0; k < T; ++k) sA[k] = r[k];
      }
      if (lane == j) {
#pragma unroll
        for (int k = 0; k < T; ++k) sB2[k] = r[k];
      }
      __syncthreads();
      const int pj = __shfl(perm, j, 64), pp = __shfl(perm, p, 64);
      if (lane == j) {
#pragma unroll
        for (int k = 0; k < T; ++k) r[k] = sA[k];
        perm = pp;
      }
      if (lane == p) {
#pragma unroll
        for (int k = 0; k < T; ++k) r[k] = sB2[k];
        perm = pj;
      }
      __syncthreads();
    }
    if (lane == j) {
#pragma unroll
      for (int k = j; k < T; ++k) sP[k] = r[k];
    }
    __syncthreads();
    if (lane > j) {
      const double l = r[j] / sP[j];
      r[j] = l;
#pragma unroll
      for (int k = j + 1; k < T; ++k) r[k] = fma(-l, sP[k], r[k]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < T; ++j) {
    A[lane * T + j] = r[j];
    sT[lane][j] = r[j];
  }
  piv[K * T + lane] = perm;
  __syncthreads();

  // L^-1 (unit lower): lane c solves L x = e_c
  {
    const int c = lane;
    double x[T];
#pragma unroll
    for (int i = 0; i < T; ++i) {
      double s = (i == c) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < i; ++k) s = fma(-sT[i][k], x[k], s);
      x[i] = (i < c) ? 0.0 : s;
    }
    double *Li = linv + K * (int64_t)T * T;
#pragma unroll
    for (int i = 0; i < T; ++i) Li[i * T + c] = x[i];
  }
  // U^-1 (upper): lane c solves U y = e_c
  {
    const int c = lane;
    double y[T];
#pragma unroll
    for (int i = T - 1; i >= 0; --i) {
      double s = (i == c) ? 1.0 : 0.0;
#pragma unroll
      for (int k = i + 1; k < T; ++k) s = fma(-sT[i][k], y[k], s);
      y[i] = (i > c) ? 0.0 : s / sT[i][i];
    }
    double *Ui = uinv + K * (int64_t)T * T;
#pragma unroll
    for (int i = 0; i < T; ++i) Ui[i * T + c] = y[i];
  }
}

// C = A B for 64x64 tiles staged in LDS; each of 256 threads owns a 4x4 block.
__device__ __forceinline__ void tile_mm(const double (*sa)[T], const double (*sb)[T],
                                        double c[4][4]) {
  const int i0 = (threadIdx.x / 16) * 4, j0 = (threadIdx.x % 16) * 4;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) c[a][b] = 0.0;
#pragma unroll 8
  for (int k = 0; k < T; ++k) {
    double av[4], bv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) av[a] = sa[i0 + a][k];
#pragma unroll
    for (int b = 0; b < 4; ++b) bv[b] = sb[k][j0 + b];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) c[a][b] = fma(av[a], bv[b], c[a][b]);
  }
}

__global__ __launch_bounds__(256) void k_panel(int64_t K, int64_t nt, int tb,
                                               double *__restrict__ band,
                                               const double *__restrict__ linv,
                                               const double *__restrict__ uinv,
                                               const int32_t *__restrict__ piv) {
  __shared__ double sa[T][T];
  __shared__ double sb[T][T];
  const int kind = blockIdx.x / tb, off = blockIdx.x % tb;
  const int32_t *pv = piv + K * T;
  if (kind == 0) {  // U_KJ = L_KK^-1 P_K A_KJ
    const int64_t J = K + 1 + off;
    if (J >= nt) return;
    double *A = tile_ptr(band, tb, K, J);
    const double *Li = linv + K * (int64_t)T * T;
    for (int k = threadIdx.x; k < T * T; k += 256) {
      sa[k / T][k % T] = Li[k];
      sb[k / T][k % T] = A[(int64_t)pv[k / T] * T + (k % T)];
    }
    __syncthreads();
    double c[4][4];
    tile_mm(sa, sb, c);
    const int i0 = (threadIdx.x / 16) * 4, j0 = (threadIdx.x % 16) * 4;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) A[(i0 + a) * T + j0 + b] = c[a][b];
  } else if (kind == 1) {  // L_IK = A_IK U_KK^-1
    const int64_t I = K + 1 + off;
    if (I >= nt) return;
    double *A = tile_ptr(band, tb, I, K);
    const double *Ui = uinv + K * (int64_t)T * T;
    for (int k = threadIdx.x; k < T * T; k += 256) {
      sa[k / T][k % T] = A[k];
      sb[k / T][k % T] = Ui[k];
    }
    __syncthreads();
    double c[4][4];
    tile_mm(sa, sb, c);
    const int i0 = (threadIdx.x / 16) * 4, j0 = (threadIdx.x % 16) * 4;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) A[(i0 + a) * T + j0 + b] = c[a][b];
  } else {  // row permutation of the L tiles left of the diagonal in row K
    const int64_t J = K - tb + off;
    if (J < 0) return;
    double *A = tile_ptr(band, tb, K, J);
    for (int k = threadIdx.x; k < T * T; k += 256) sa[k / T][k % T] = A[k];
    __syncthreads();
    for (int k = threadIdx.x; k < T * T; k += 256) A[k] = sa[pv[k / T]][k % T];
  }
}

__global__ __launch_bounds__(256) void k_update(int64_t K, int64_t nt, int tb,
                                                double *__restrict__ band) {
  __shared__ double sa[T][T];
  __shared__ double sb[T][T];
  const int64_t I = K + 1 + blockIdx.x / tb, J = K + 1 + blockIdx.x % tb;
  if (I >= nt || J >= nt) return;
  const double *Lt = tile_ptr(band, tb, I, K);
  const double *Ut = tile_ptr(band, tb, K, J);
  for (int k = threadIdx.x; k < T * T; k += 256) {
    sa[k / T][k % T] = Lt[k];
    sb[k / T][k % T] = Ut[k];
  }
  __syncthreads();
  double c[4][4];
  tile_mm(sa, sb, c);
  double *A = tile_ptr(band, tb, I, J);
  const int i0 = (threadIdx.x / 16) * 4, j0 = (threadIdx.x % 16) * 4;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) A[(i0 + a) * T + j0 + b] -= c[a][b];
}

// ------------------------------------------------------------------ solves

// t = P_I b_I - sum_K L_IK y_K ;  y_I = L_II^-1 t
__global__ __launch_bounds__(256) void k_fwd(int64_t I, int tb, const double *__restrict__ band,
                                             const double *__restrict__ linv,
                                             const int32_t *__restrict__ piv,
                                             const double *__restrict__ b,
                                             double *__restrict__ y) {
  __shared__ double st[T];
  const int i = threadIdx.x / 4, q = threadIdx.x % 4;
  double s = 0.0;
  const int64_t K0 = I - tb < 0 ? 0 : I - tb;
  for (int64_t K = K0; K < I; ++K) {
    const double *L = tile_ptr(const_cast<double *>(band), tb, I, K) + i * T + q * 16;
    const double *yk = y + K * T + q * 16;
#pragma unroll
    for (int k = 0; k < 16; ++k) s = fma(L[k], yk[k], s);
  }
  s += __shfl_xor(s, 1, 64);
  s += __shfl_xor(s, 2, 64);
  if (q == 0) st[i] = b[I * T + piv[I * T + i]] - s;
  __syncthreads();
  const double *Li = linv + I * (int64_t)T * T + i * T + q * 16;
  double u = 0.0;
#pragma unroll
  for (int k = 0; k < 16; ++k) u = fma(Li[k], st[q * 16 + k], u);
  u += __shfl_xor(u, 1, 64);
  u += __shfl_xor(u, 2, 64);
  if (q == 0) y[I * T + i] = u;
}

// t = y_I - sum_J U_IJ x_J ;  x_I = U_II^-1 t   (x may alias y)
__global__ __launch_bounds__(256) void k_bwd(int64_t I, int64_t nt, int tb,
                                             const double *__restrict__ band,
                                             const double *__restrict__ uinv,
                                             double *__restrict__ x) {
  __shared__ double st[T];
  const int i = threadIdx.x / 4, q = threadIdx.x % 4;
  double s = 0.0;
  const int64_t J1 = I + tb > nt - 1 ? nt - 1 : I + tb;
  for (int64_t J = I + 1; J <= J1; ++J) {
    const double *U = tile_ptr(const_cast<double *>(band), tb, I, J) + i * T + q * 16;
    const double *xj = x + J * T + q * 16;
#pragma unroll
    for (int k = 0; k < 16; ++k) s = fma(U[k], xj[k], s);
  }
  s += __shfl_xor(s, 1, 64);
  s += __shfl_xor(s, 2, 64);
  if (q == 0) st[i] = x[I * T + i] - s;
  __syncthreads();
  const double *Ui = uinv + I * (int64_t)T * T + i * T + q * 16;
  double u = 0.0;
#pragma unroll
  for (int k = 0; k < 16; ++k) u = fma(Ui[k], st[q * 16 + k], u);
  u += __shfl_xor(u, 1, 64);
  u += __shfl_xor(u, 2, 64);
  __syncthreads();
  if (q == 0) x[I * T + i] = u;
}

// ------------------------------------------------------------ BSR and vectors

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

void band_fill(Ctx &c) {
  const int64_t nband = c.nt * (2 * c.tb + 1) * (int64_t)T * T;
  hipLaunchKernelGGL(k_zero, dim3(nblk(nband)), dim3(256), 0, c.stream, nband, c.d_band);
  const int64_t nb = c.nnzb * c.ndf * c.ndf;
  hipLaunchKernelGGL(k_bsr_to_band, dim3(nblk(nb)), dim3(256), 0, c.stream, c.nnzb, c.ndf, c.tb,
                     c.d_brow, c.d_colidx, c.d_bsr, c.d_band);
  const int64_t npad = c.nt * T;
  if (npad > c.nx)
    hipLaunchKernelGGL(k_band_pad, dim3((unsigned)((npad - c.nx + 63) / 64)), dim3(64), 0, c.stream,
                       c.nx, npad, c.tb, c.d_band);
  MAFEM_HIP(hipGetLastError());
}

void band_factor(Ctx &c) {
  for (int64_t K = 0; K < c.nt; ++K) {
    hipLaunchKernelGGL(k_diag, dim3(1), dim3(64), 0, c.stream, K, c.tb, c.d_band, c.d_linv,
                       c.d_uinv, c.d_piv);
    hipLaunchKernelGGL(k_panel, dim3(3 * c.tb), dim3(256), 0, c.stream, K, c.nt, c.tb, c.d_band,
                       c.d_linv, c.d_uinv, c.d_piv);
    hipLaunchKernelGGL(k_update, dim3(c.tb * c.tb), dim3(256), 0, c.stream, K, c.nt, c.tb,
                       c.d_band);
  }
  MAFEM_HIP(hipGetLastError());
}

void band_solve(Ctx &c, const double *d_rhs, double *d_x) {
  for (int64_t I = 0; I < c.nt; ++I)
    hipLaunchKernelGGL(k_fwd, dim3(1), dim3(256), 0, c.stream, I, c.tb, c.d_band, c.d_linv,
                       c.d_piv, d_rhs, d_x);
  for (int64_t I = c.nt - 1; I >= 0; --I)
    hipLaunchKernelGGL(k_bwd, dim3(1), dim3(256), 0, c.stream, I, c.nt, c.tb, c.d_band, c.d_uinv,
                       d_x);
  MAFEM_HIP(hipGetLastError());
}

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

}  // namespace mafem
