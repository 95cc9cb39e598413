// Banded LU of the extended node-block system on gfx950, stored as dense
// 64x64 tiles inside the band, in FP64 or (mixed-precision path) FP32.
//
// Natural node numbering (i1 + num1np i2) bounds the coupling of two nodes
// that share an element to |Δnode| <= 2 num1np + 2, so the scalar matrix has
// half bandwidth kl = (2 num1np + 3) ndf - 1 and LU without pivoting across
// tile rows keeps the fill inside that band.  Partial pivoting is applied
// inside each diagonal tile.  Per tile step K:
//   k_diag   : P_K A_KK = L_KK U_KK in one workgroup (tile in LDS), then
//              explicit inverses of L_KK and U_KK;
//   k_panel  : U_KJ = L_KK^-1 P_K A_KJ, L_IK = A_IK U_KK^-1, and P_K applied to
//              the L tiles already stored in tile row K;
//   k_update : A_IJ -= L_IK U_KJ over the (tb x tb) trailing window.
// The triangular sweeps run in one workgroup each (the tile rows form a
// dependency chain) and accumulate in FP64 whatever the storage precision.
// This replaces `K_gl \ r_gl` (src/analysis/FiniteElement.jl:38, UMFPACK).
#include <algorithm>

#include "internal.h"

namespace mafem {

constexpr int T = TILE;

template <class S>
__device__ __forceinline__ S *tile_ptr(S *band, int tb, int64_t I, int64_t J) {
  return band + ((I * (2 * tb + 1) + (J - I + tb)) * (int64_t)T * T);
}

// ----------------------------------------------------------------- band fill

template <class S>
__global__ __launch_bounds__(256) void k_zero_t(int64_t n, S *__restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) x[t] = S(0);
}

template <class S>
__global__ __launch_bounds__(256) void k_bsr_to_band(int64_t nnzb, int ndf, int tb,
                                                     const int32_t *__restrict__ brow,
                                                     const int32_t *__restrict__ colidx,
                                                     const double *__restrict__ bsr,
                                                     S *__restrict__ band) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int bb = ndf * ndf;
  if (t >= nnzb * bb) return;
  const int64_t blk = t / bb;
  const int rc = (int)(t % bb), r = rc / ndf, c = rc % ndf;
  const int64_t R = (int64_t)brow[blk] * ndf + r, C = (int64_t)colidx[blk] * ndf + c;
  tile_ptr(band, tb, R / T, C / T)[(R % T) * T + (C % T)] = (S)bsr[t];
}

// identity on the padding rows beyond the extended system
template <class S>
__global__ void k_band_pad(int64_t nx, int64_t npad, int tb, S *__restrict__ band) {
  const int64_t R = nx + (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (R >= npad) return;
  tile_ptr(band, tb, R / T, R / T)[(R % T) * T + (R % T)] = S(1);
}

// ------------------------------------------------------------ factorization

template <class S>
__device__ __forceinline__ void wave_argmax(S &v, int &idx) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const S ov = __shfl_xor(v, off, 64);
    const int oi = __shfl_xor(idx, off, 64);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
}

// Makes this wave's LDS writes visible to its other lanes before later reads.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Factors the 64x64 diagonal tile with partial pivoting (P A = L U) and forms
// L^-1 and U^-1.  The tile is held column-major in LDS (sT[col][row], padded)
// so that lane = row accesses are contiguous.  Wave 0 picks the pivot, swaps
// rows and forms the multipliers; all four waves share the rank-1 update by
// columns.  Then wave 0 inverts L and wave 1 inverts U, both in place (the
// strictly lower and the upper parts are disjoint).
template <class S>
__global__ __launch_bounds__(256) void k_diag(int64_t K, int tb, S *__restrict__ band,
                                              S *__restrict__ linv, S *__restrict__ uinv,
                                              int32_t *__restrict__ piv) {
  __shared__ S sT[T][T + 1];  // sT[col][row]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  S *A = tile_ptr(band, tb, K, K);
  for (int k = tid; k < T * T; k += 256) sT[k % T][k / T] = A[k];
  int perm = lane;
  __syncthreads();
  for (int j = 0; j < T; ++j) {
    if (wv == 0) {
      S v = (lane >= j) ? fabs(sT[j][lane]) : S(-1);
      int p = lane;
      wave_argmax(v, p);
      if (p != j) {  // swap rows j and p: lane k swaps column k
        const S a = sT[lane][j], b = sT[lane][p];
        sT[lane][j] = b;
        sT[lane][p] = a;
        const int pj = __shfl(perm, j, 64), pp = __shfl(perm, p, 64);
        if (lane == j) perm = pp;
        if (lane == p) perm = pj;
      }
      wave_lds_sync();
      if (lane > j) sT[j][lane] = sT[j][lane] / sT[j][j];
    }
    __syncthreads();
    if (lane > j) {
      const S l = sT[j][lane];
#pragma unroll 4
      for (int k = j + 1 + wv; k < T; k += 4) sT[k][lane] = fma(-l, sT[k][j], sT[k][lane]);
    }
    __syncthreads();
  }
  for (int k = tid; k < T * T; k += 256) A[k] = sT[k % T][k / T];
  if (wv == 0) piv[K * T + lane] = perm;
  __syncthreads();  // the tile is copied out before the in-place inversions

  if (wv == 0) {
    // V = L^-1 (unit lower), columns j descending: V_ij = -Σ_{k=j+1}^{i} V_ik L_kj
    const int i = lane;
    for (int j = T - 2; j >= 0; --j) {
      S s0 = 0, s1 = 0;
      if (i > j) {
        int k = j + 1;
        for (; k + 1 < i; k += 2) {
          s0 = fma(sT[k][i], sT[j][k], s0);
          s1 = fma(sT[k + 1][i], sT[j][k + 1], s1);
        }
        if (k < i) s0 = fma(sT[k][i], sT[j][k], s0);
        s0 += sT[j][i];  // k = i: V_ii = 1 times L_ij
      }
      wave_lds_sync();
      if (i > j) sT[j][i] = -(s0 + s1);
      wave_lds_sync();
    }
  } else if (wv == 1) {
    // W = U^-1 (upper), columns j ascending: W_ij = -Σ_{k=i}^{j-1} W_ik U_kj / U_jj
    const int i = lane;
    for (int j = 0; j < T; ++j) {
      const S ujj = sT[j][j];
      S s0 = 0, s1 = 0;
      if (i < j) {
        int k = i;
        for (; k + 1 < j; k += 2) {
          s0 = fma(sT[k][i], sT[j][k], s0);
          s1 = fma(sT[k + 1][i], sT[j][k + 1], s1);
        }
        if (k < j) s0 = fma(sT[k][i], sT[j][k], s0);
      }
      wave_lds_sync();
      if (i < j) sT[j][i] = -(s0 + s1) / ujj;
      if (i == j) sT[j][j] = S(1) / ujj;
      wave_lds_sync();
    }
  }
  __syncthreads();
  S *Li = linv + K * (int64_t)T * T;
  S *Ui = uinv + K * (int64_t)T * T;
  for (int k = tid; k < T * T; k += 256) {
    const int i = k / T, jj = k % T;
    Li[k] = (jj < i) ? sT[jj][i] : (jj == i ? S(1) : S(0));
    Ui[k] = (jj >= i) ? sT[jj][i] : S(0);
  }
}

// C = A B for 64x64 tiles staged in LDS; each of 256 threads owns a 4x4 block.
template <class S>
__device__ __forceinline__ void tile_mm(const S (*sa)[T], const S (*sb)[T], S c[4][4]) {
  const int i0 = (threadIdx.x / 16) * 4, j0 = (threadIdx.x % 16) * 4;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) c[a][b] = 0;
#pragma unroll 8
  for (int k = 0; k < T; ++k) {
    S av[4], bv[4];
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

template <class S>
__global__ __launch_bounds__(256) void k_panel(int64_t K, int64_t nt, int tb, S *__restrict__ band,
                                               const S *__restrict__ linv,
                                               const S *__restrict__ uinv,
                                               const int32_t *__restrict__ piv) {
  __shared__ S sa[T][T];
  __shared__ S sb[T][T];
  const int kind = blockIdx.x / tb, off = blockIdx.x % tb;
  const int32_t *pv = piv + K * T;
  const int i0 = (threadIdx.x / 16) * 4, j0 = (threadIdx.x % 16) * 4;
  if (kind == 0) {  // U_KJ = L_KK^-1 P_K A_KJ
    const int64_t J = K + 1 + off;
    if (J >= nt) return;
    S *A = tile_ptr(band, tb, K, J);
    const S *Li = linv + K * (int64_t)T * T;
    for (int k = threadIdx.x; k < T * T; k += 256) {
      sa[k / T][k % T] = Li[k];
      sb[k / T][k % T] = A[(int64_t)pv[k / T] * T + (k % T)];
    }
    __syncthreads();
    S c[4][4];
    tile_mm<S>(sa, sb, c);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) A[(i0 + a) * T + j0 + b] = c[a][b];
  } else if (kind == 1) {  // L_IK = A_IK U_KK^-1
    const int64_t I = K + 1 + off;
    if (I >= nt) return;
    S *A = tile_ptr(band, tb, I, K);
    const S *Ui = uinv + K * (int64_t)T * T;
    for (int k = threadIdx.x; k < T * T; k += 256) {
      sa[k / T][k % T] = A[k];
      sb[k / T][k % T] = Ui[k];
    }
    __syncthreads();
    S c[4][4];
    tile_mm<S>(sa, sb, c);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) A[(i0 + a) * T + j0 + b] = c[a][b];
  } else {  // row permutation of the L tiles left of the diagonal in row K
    const int64_t J = K - tb + off;
    if (J < 0) return;
    S *A = tile_ptr(band, tb, K, J);
    for (int k = threadIdx.x; k < T * T; k += 256) sa[k / T][k % T] = A[k];
    __syncthreads();
    for (int k = threadIdx.x; k < T * T; k += 256) A[k] = sa[pv[k / T]][k % T];
  }
}

template <class S>
__global__ __launch_bounds__(256) void k_update(int64_t K, int64_t nt, int tb, S *__restrict__ band) {
  __shared__ S sa[T][T];
  __shared__ S sb[T][T];
  const int64_t I = K + 1 + blockIdx.x / tb, J = K + 1 + blockIdx.x % tb;
  if (I >= nt || J >= nt) return;
  const S *Lt = tile_ptr(band, tb, I, K);
  const S *Ut = tile_ptr(band, tb, K, J);
  for (int k = threadIdx.x; k < T * T; k += 256) {
    sa[k / T][k % T] = Lt[k];
    sb[k / T][k % T] = Ut[k];
  }
  __syncthreads();
  S c[4][4];
  tile_mm<S>(sa, sb, c);
  S *A = tile_ptr(band, tb, I, J);
  const int i0 = (threadIdx.x / 16) * 4, j0 = (threadIdx.x % 16) * 4;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) A[(i0 + a) * T + j0 + b] -= c[a][b];
}

// ------------------------------------------------------------ triangular solves
//
// The band's tile rows form a dependency chain, so each sweep runs in one
// 1024-thread workgroup: no inter-workgroup synchronisation, the last tb
// solution tiles stay in an LDS ring, and each step streams the <= tb band
// tiles of its row with 16 threads per tile row (4 columns each).

constexpr int SW = 1024;          // threads of a sweep workgroup
constexpr int RING = SOLVE_RING;  // LDS ring of solution tiles (>= tb + 1)

__device__ __forceinline__ double row16_sum(double s) {
  s += __shfl_xor(s, 1, 64);
  s += __shfl_xor(s, 2, 64);
  s += __shfl_xor(s, 4, 64);
  s += __shfl_xor(s, 8, 64);
  return s;
}

template <class S>
struct Vec4;
template <>
struct Vec4<double> {
  using type = double4;
};
template <>
struct Vec4<float> {
  using type = float4;
};

// thread (i = tid / 16, q = tid % 16) covers columns 4q..4q+3 of row i
template <class S>
__device__ __forceinline__ double tile_row4(const S *M, const double *v) {
  const int i = threadIdx.x >> 4, q = threadIdx.x & 15;
  const typename Vec4<S>::type m =
      *reinterpret_cast<const typename Vec4<S>::type *>(M + i * T + 4 * q);
  const double *x = v + 4 * q;
  return fma((double)m.x, x[0], fma((double)m.y, x[1], fma((double)m.z, x[2], (double)m.w * x[3])));
}

template <class S>
__global__ __launch_bounds__(SW) void k_fwd_sweep(int64_t nt, int tb, const S *__restrict__ band,
                                                  const S *__restrict__ linv,
                                                  const int32_t *__restrict__ piv,
                                                  const double *__restrict__ b,
                                                  double *__restrict__ y) {
  __shared__ double ring[RING][T];
  __shared__ double st[T];
  const int i = threadIdx.x >> 4, q = threadIdx.x & 15;
  for (int64_t I = 0; I < nt; ++I) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    const int64_t K0 = I - tb < 0 ? 0 : I - tb;
    int64_t K = K0;
    for (; K + 3 < I; K += 4) {
      s0 += tile_row4(tile_ptr(band, tb, I, K), ring[K % RING]);
      s1 += tile_row4(tile_ptr(band, tb, I, K + 1), ring[(K + 1) % RING]);
      s2 += tile_row4(tile_ptr(band, tb, I, K + 2), ring[(K + 2) % RING]);
      s3 += tile_row4(tile_ptr(band, tb, I, K + 3), ring[(K + 3) % RING]);
    }
    for (; K < I; ++K) s0 += tile_row4(tile_ptr(band, tb, I, K), ring[K % RING]);
    const double s = row16_sum((s0 + s1) + (s2 + s3));
    if (q == 0) st[i] = b[I * T + piv[I * T + i]] - s;
    __syncthreads();
    const double u = row16_sum(tile_row4(linv + I * (int64_t)T * T, st));
    if (q == 0) {
      ring[I % RING][i] = u;
      y[I * T + i] = u;
    }
    __syncthreads();
  }
}

template <class S>
__global__ __launch_bounds__(SW) void k_bwd_sweep(int64_t nt, int tb, const S *__restrict__ band,
                                                  const S *__restrict__ uinv,
                                                  double *__restrict__ x) {
  __shared__ double ring[RING][T];
  __shared__ double st[T];
  const int i = threadIdx.x >> 4, q = threadIdx.x & 15;
  for (int64_t I = nt - 1; I >= 0; --I) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    const int64_t J1 = I + tb > nt - 1 ? nt - 1 : I + tb;
    int64_t J = I + 1;
    for (; J + 3 <= J1; J += 4) {
      s0 += tile_row4(tile_ptr(band, tb, I, J), ring[J % RING]);
      s1 += tile_row4(tile_ptr(band, tb, I, J + 1), ring[(J + 1) % RING]);
      s2 += tile_row4(tile_ptr(band, tb, I, J + 2), ring[(J + 2) % RING]);
      s3 += tile_row4(tile_ptr(band, tb, I, J + 3), ring[(J + 3) % RING]);
    }
    for (; J <= J1; ++J) s0 += tile_row4(tile_ptr(band, tb, I, J), ring[J % RING]);
    const double s = row16_sum((s0 + s1) + (s2 + s3));
    if (q == 0) st[i] = x[I * T + i] - s;
    __syncthreads();
    const double u = row16_sum(tile_row4(uinv + I * (int64_t)T * T, st));
    if (q == 0) {
      ring[I % RING][i] = u;
      x[I * T + i] = u;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ host side

static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

template <class S>
static void fill_t(Ctx &c, S *band) {
  const int64_t nband = c.nt * (2 * c.tb + 1) * (int64_t)T * T;
  hipLaunchKernelGGL(k_zero_t<S>, dim3(nblk(nband)), dim3(256), 0, c.stream, nband, band);
  const int64_t nb = c.nnzb * c.ndf * c.ndf;
  hipLaunchKernelGGL(k_bsr_to_band<S>, dim3(nblk(nb)), dim3(256), 0, c.stream, c.nnzb, c.ndf, c.tb,
                     c.d_brow, c.d_colidx, c.d_bsr, band);
  const int64_t npad = c.nt * T;
  if (npad > c.nx)
    hipLaunchKernelGGL(k_band_pad<S>, dim3((unsigned)((npad - c.nx + 63) / 64)), dim3(64), 0,
                       c.stream, c.nx, npad, c.tb, band);
  MAFEM_HIP(hipGetLastError());
}

template <class S>
static void factor_t(Ctx &c, S *band, S *li, S *ui) {
  for (int64_t K = 0; K < c.nt; ++K) {
    hipLaunchKernelGGL(k_diag<S>, dim3(1), dim3(256), 0, c.stream, K, c.tb, band, li, ui, c.d_piv);
    hipLaunchKernelGGL(k_panel<S>, dim3(3 * c.tb), dim3(256), 0, c.stream, K, c.nt, c.tb, band, li,
                       ui, c.d_piv);
    hipLaunchKernelGGL(k_update<S>, dim3(c.tb * c.tb), dim3(256), 0, c.stream, K, c.nt, c.tb, band);
  }
  MAFEM_HIP(hipGetLastError());
}

template <class S>
static void solve_t(Ctx &c, const S *band, const S *li, const S *ui, const double *d_rhs,
                    double *d_x) {
  hipLaunchKernelGGL(k_fwd_sweep<S>, dim3(1), dim3(SW), 0, c.stream, c.nt, c.tb, band, li, c.d_piv,
                     d_rhs, d_x);
  hipLaunchKernelGGL(k_bwd_sweep<S>, dim3(1), dim3(SW), 0, c.stream, c.nt, c.tb, band, ui, d_x);
  MAFEM_HIP(hipGetLastError());
}

// Allocates the band storage of the requested precision on first use.
static void ensure_band(Ctx &c, int prec) {
  const int64_t ntile = c.nt * (2 * c.tb + 1) * (int64_t)T * T, ninv = c.nt * (int64_t)T * T;
  const size_t es = (prec == 32) ? sizeof(float) : sizeof(double);
  if (c.band_prec == prec && c.d_band) return;
  if (c.d_band) {
    MAFEM_HIP(hipStreamSynchronize(c.stream));
    (void)hipFree(c.d_band);
    (void)hipFree(c.d_linv);
    (void)hipFree(c.d_uinv);
    c.d_band = c.d_linv = c.d_uinv = nullptr;
  }
  void *p = nullptr;
  if (hipMalloc(&p, es * ntile) != hipSuccess) {
    (void)hipGetLastError();
    throw HipError(hipErrorOutOfMemory, "band allocation", __FILE__, __LINE__);
  }
  c.d_band = p;
  MAFEM_HIP(hipMalloc(&p, es * ninv));
  c.d_linv = p;
  MAFEM_HIP(hipMalloc(&p, es * ninv));
  c.d_uinv = p;
  c.band_prec = prec;
  c.band_bytes = (int64_t)(es * ntile);
}

void band_fill(Ctx &c, int prec) {
  ensure_band(c, prec);
  if (prec == 32)
    fill_t<float>(c, (float *)c.d_band);
  else
    fill_t<double>(c, (double *)c.d_band);
}

void band_factor(Ctx &c) {
  ++c.n_factor;
  if (c.band_prec == 32)
    factor_t<float>(c, (float *)c.d_band, (float *)c.d_linv, (float *)c.d_uinv);
  else
    factor_t<double>(c, (double *)c.d_band, (double *)c.d_linv, (double *)c.d_uinv);
}

void band_solve(Ctx &c, const double *d_rhs, double *d_x) {
  ++c.n_band_solve;
  if (c.band_prec == 32)
    solve_t<float>(c, (const float *)c.d_band, (const float *)c.d_linv, (const float *)c.d_uinv,
                   d_rhs, d_x);
  else
    solve_t<double>(c, (const double *)c.d_band, (const double *)c.d_linv,
                    (const double *)c.d_uinv, d_rhs, d_x);
}

}  // namespace mafem
