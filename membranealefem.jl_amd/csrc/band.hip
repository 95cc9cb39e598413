// Banded LU of the extended node-block system on gfx950, stored as dense
// 64x64 tiles inside the band, in FP64 or (mixed-precision path) FP32.
//
// Natural node numbering (i1 + num1np i2) bounds the coupling of two nodes
// that share an element to |Δnode| <= 2 num1np + 2, so the scalar matrix has
// half bandwidth kl = (2 num1np + 3) ndf - 1 and LU without pivoting across
// tile rows keeps the fill inside that band.  Partial pivoting is applied
// inside each diagonal tile.  Per tile step K:
//   k_diag_mm: P_K A_KK = L_KK U_KK by one block (pivot steps in one
//              wavefront, rank-16 trailing updates on the matrix cores);
//   k_panel  : U_KJ = L_KK^-1 P_K A_KJ, L_IK = A_IK U_KK^-1 (triangular
//              solves, one wavefront per tile) and P_K applied to the L tiles
//              already stored in tile row K;
//   k_update : A_IJ -= L_IK U_KJ over the (tb x tb) trailing window.
// Afterwards a fully parallel pass inverts the diagonal tiles and folds the
// inverses into the off-diagonal tiles.  The triangular sweeps run in one
// workgroup each (the tile rows form a dependency chain) and accumulate in
// FP64 whatever the storage precision.
// This replaces `K_gl \ r_gl` (src/analysis/FiniteElement.jl:38, UMFPACK).
#include <algorithm>

#include "internal.h"

namespace mafem {

constexpr int T = TILE;

template <class S>
__device__ __forceinline__ S *tile_ptr(S *band, int tb, int64_t I, int64_t J) {
  return band + ((I * (2 * tb + 1) + (J - I + tb)) * (int64_t)T * T);
}

// Values published between workgroups of one launch (triangular sweeps,
// factorization look-ahead) go through words pre-filled with this pattern (a
// signalling NaN that no result takes), written and read with agent-scope
// atomics only; a word is its own ready flag.  Waits are bounded by SPIN_LIMIT.
constexpr unsigned long long SENTINEL = 0x7FF4DEADBEEF1234ULL;
constexpr long SPIN_LIMIT = 20000000L;
constexpr int PUBW = T * T + T;  // look-ahead staging: a factored tile, then its pivots

template <class S> __device__ __forceinline__ unsigned long long pub_bits(S x);
template <> __device__ __forceinline__ unsigned long long pub_bits<double>(double x) {
  return (unsigned long long)__double_as_longlong(x);
}
template <> __device__ __forceinline__ unsigned long long pub_bits<float>(float x) {
  return (unsigned long long)__float_as_uint(x);
}
template <class S> __device__ __forceinline__ S pub_val(unsigned long long v);
template <> __device__ __forceinline__ double pub_val<double>(unsigned long long v) {
  return __longlong_as_double((long long)v);
}
template <> __device__ __forceinline__ float pub_val<float>(unsigned long long v) {
  return __uint_as_float((unsigned)v);
}

// ----------------------------------------------------------------- band fill

template <class S>
__global__ __launch_bounds__(256) void k_zero_t(int64_t n, S *__restrict__ x) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
    x[t] = S(0);
}

template <class S>
__device__ __forceinline__ void bsr_to_band_body(int64_t bid, int64_t nnzb, int ndf, int tb,
                                                     const int32_t *__restrict__ brow,
                                                     const int32_t *__restrict__ colidx,
                                                     const double *__restrict__ bsr,
                                                     S *__restrict__ band) {
  const int64_t t = (int64_t)bid * 256 + threadIdx.x;
  const int bb = ndf * ndf;
  if (t >= nnzb * bb) return;
  const int64_t blk = t / bb;
  const int rc = (int)(t % bb), r = rc / ndf, c = rc % ndf;
  const int64_t R = (int64_t)brow[blk] * ndf + r, C = (int64_t)colidx[blk] * ndf + c;
  tile_ptr(band, tb, R / T, C / T)[(R % T) * T + (C % T)] = (S)bsr[t];
}

template <class S>
__global__ __launch_bounds__(256) void k_bsr_to_band(int64_t nnzb, int ndf, int tb,
                                                     const int32_t *__restrict__ brow,
                                                     const int32_t *__restrict__ colidx,
                                                     const double *__restrict__ bsr,
                                                     S *__restrict__ band) {
  bsr_to_band_body<S>((int64_t)blockIdx.x, nnzb, ndf, tb, brow, colidx, bsr, band);
}

// identity on the padding rows beyond the extended system
template <class S>
__global__ void k_band_pad(int64_t nx, int64_t npad, int tb, S *__restrict__ band) {
  const int64_t R = nx + (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (R >= npad) return;
  tile_ptr(band, tb, R / T, R / T)[(R % T) * T + (R % T)] = S(1);
}

// ------------------------------------------- two-ended layout (Ctx::tw_a > 0)
//
// Scalar row R belongs to the top band BT while R < ns = (a + w) TILE (T and
// the separator S, original order); the bottom band BR holds B and S in
// reversed order, R -> N - 1 - R (N = nt TILE).  An entry between two rows
// of S goes to BT only: BR's separator block starts at zero and collects B's
// Schur complement.  T and B never couple (they are more than tb tiles apart).
template <class S>
__global__ __launch_bounds__(256) void k_bsr_to_band_tw(int64_t nnzb, int ndf, int tb, int64_t ns,
                                                        int64_t N, const int32_t *__restrict__ brow,
                                                        const int32_t *__restrict__ colidx,
                                                        const double *__restrict__ bsr,
                                                        S *__restrict__ bt, S *__restrict__ br) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int bb = ndf * ndf;
  if (t >= nnzb * bb) return;
  const int64_t blk = t / bb;
  const int rc = (int)(t % bb), r = rc / ndf, c = rc % ndf;
  int64_t R = (int64_t)brow[blk] * ndf + r, C = (int64_t)colidx[blk] * ndf + c;
  S *band = bt;
  if (R >= ns || C >= ns) {
    R = N - 1 - R;
    C = N - 1 - C;
    band = br;
  }
  tile_ptr(band, tb, R / T, C / T)[(R % T) * T + (C % T)] = (S)bsr[t];
}

// identity on the padding rows [nx, N), which lie in B (reversed: BR's first rows)
template <class S>
__global__ void k_band_pad_tw(int64_t nx, int64_t N, int tb, S *__restrict__ br) {
  const int64_t R = nx + (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (R >= N) return;
  const int64_t Rr = N - 1 - R;
  tile_ptr(br, tb, Rr / T, Rr / T)[(Rr % T) * T + (Rr % T)] = S(1);
}

// BT's separator tile (I, J) += BR's tile (nt-1-I, nt-1-J) read in reversed
// element order (w x w tiles, one block each)
template <class S>
__global__ __launch_bounds__(256) void k_sep_add(int64_t a, int w, int tb, int64_t nt,
                                                 S *__restrict__ bt, const S *__restrict__ br) {
  const int64_t I = a + blockIdx.x / w, J = a + blockIdx.x % w;
  S *d = tile_ptr(bt, tb, I, J);
  const S *s = tile_ptr(br, tb, nt - 1 - I, nt - 1 - J);
  for (int e = threadIdx.x; e < T * T; e += 256) d[e] += s[T * T - 1 - e];
}

// out_I -= Σ_K M_IK v_K over the band tiles (I, K), K in [max(klo, I - tb),
// min(khi, I + tb + 1)), for rows I = I0 + blockIdx.x; tiles in increasing K.
// vrev / orev > 0 address v / out in reversed order (element k of tile K at
// vrev - 1 - (K TILE + k)).  ATOMIC: v is a multi-workgroup sweep output.
template <class S, bool ATOMIC>
__global__ __launch_bounds__(256) void k_rows_sub(int tb, const S *__restrict__ band, int64_t I0,
                                                  int64_t klo, int64_t khi,
                                                  const double *__restrict__ v, int64_t vrev,
                                                  double *__restrict__ out, int64_t orev) {
  __shared__ double sv[T];
  const int64_t I = I0 + blockIdx.x;
  const int i = threadIdx.x / 4, q = threadIdx.x % 4;
  const int64_t K0 = I - tb > klo ? I - tb : klo;
  const int64_t K1 = I + tb + 1 < khi ? I + tb + 1 : khi;
  double s = 0.0;
  for (int64_t K = K0; K < K1; ++K) {
    __syncthreads();
    if (threadIdx.x < T) {
      int64_t k = K * T + threadIdx.x;
      if (vrev > 0) k = vrev - 1 - k;
      sv[threadIdx.x] = ATOMIC ? __longlong_as_double((long long)__hip_atomic_load(
                                     reinterpret_cast<const unsigned long long *>(v) + k,
                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                               : v[k];
    }
    __syncthreads();
    const S *m = tile_ptr(band, tb, I, K) + i * T + q * 16;
#pragma unroll
    for (int k = 0; k < 16; ++k) s = fma((double)m[k], sv[q * 16 + k], s);
  }
  s += __shfl_xor(s, 1, 64);
  s += __shfl_xor(s, 2, 64);
  if (q == 0) {
    int64_t k = I * T + i;
    if (orev > 0) k = orev - 1 - k;
    out[k] -= s;
  }
}

// dst[k] = src[N - 1 - k] for k in [k0, k0 + n); ATOMIC: src is a sweep output
template <bool ATOMIC>
__global__ __launch_bounds__(256) void k_rev_copy(int64_t k0, int64_t n, int64_t N,
                                                  const double *__restrict__ src,
                                                  double *__restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const int64_t k = k0 + t;
  dst[k] = ATOMIC ? __longlong_as_double((long long)__hip_atomic_load(
                        reinterpret_cast<const unsigned long long *>(src) + (N - 1 - k),
                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                  : src[N - 1 - k];
}

// ------------------------------------------------------------ factorization

// Maximum over the wavefront of a 32-bit key, delivered as a scalar: DPP
// row-local steps, then row broadcasts that leave the total in lane 63.
// max over the wavefront.  The moved-in operand is 0 (the identity of the
// unsigned max) where a row is masked off, which lets the compiler fold each
// move into v_max_u32_dpp: one instruction per stage.
__device__ __forceinline__ unsigned wave_max_u32(unsigned x) {
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, true));
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, true));
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, true));
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x140, 0xF, 0xF, true));
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, true));
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, true));
  return (unsigned)__builtin_amdgcn_readlane((int)x, 63);
}

// Pivot key of a candidate row: |a| rounded to FP32 (order-preserving as an
// unsigned integer) with the low 6 bits replaced by 63 - lane, so the largest
// magnitude wins and near-ties (relative 2^-17) go to the lowest row index.
// Rows already chosen carry key 0 and never win.
template <class S>
__device__ __forceinline__ unsigned pivot_key(S a, bool candidate, int lane) {
  unsigned fb = __float_as_uint((float)fabs(a));
  fb = fb < 64u ? 64u : fb;
  return candidate ? ((fb & ~63u) | (unsigned)(63 - lane)) : 0u;
}

// value of x held by lane l (l wave-uniform), delivered as a scalar
__device__ __forceinline__ double lane_bcast(double x, int l) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float lane_bcast(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// amax[0] = max |a| over n values (non-negative doubles order like their bits)
template <class S>
__global__ __launch_bounds__(256) void k_absmax(int64_t n, const S *__restrict__ a,
                                                unsigned long long *__restrict__ amax) {
  double m = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
    m = fmax(m, fabs((double)a[t]));
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0 && m > 0.0)
    atomicMax(amax, (unsigned long long)__double_as_longlong(m));
}

__global__ void k_zero_u64(unsigned long long *__restrict__ p) {
  if (threadIdx.x == 0) p[0] = 0;
}

// Triangular solves of a panel tile by one wavefront, both tiles in LDS (the
// factors read with wave-uniform addresses, rows unpadded so that they load
// 16 bytes at a time; the tile's rows padded against bank conflicts):
// kind 0, X <- L^-1 X (unit lower L of the factored diagonal tile D; lane =
// column, forward substitution); kind 1, X <- X U^-1 (lane = row, rdiag[k] =
// 1 / D_kk).  The tile is processed in blocks of 8 unknowns by loops whose
// bodies are small (a fully unrolled 64 x 64 substitution is ~9 000
// straight-line instructions, and fetching them dominated the kernel).  Every
// entry receives its updates in increasing order of the eliminated index, as
// in the unblocked substitution.
template <class S>
__device__ __forceinline__ void trsm_wave(int kind, const S (*sD)[T], S (*sX)[T + 1],
                                          const S *rdiag) {
  constexpr int B = 8;
  const int lane = (int)(threadIdx.x & 63);
  if (kind == 0) {
    // x_i -= D_ik x_k for k < i, k increasing (unit lower L)
#pragma unroll 1
    for (int b = 0; b < T / B; ++b) {
      S xb[B];
#pragma unroll
      for (int r = 0; r < B; ++r) xb[r] = sX[B * b + r][lane];
#pragma unroll 1
      for (int c = 0; c < b; ++c) {
        S xc[B];
#pragma unroll
        for (int q = 0; q < B; ++q) xc[q] = sX[B * c + q][lane];
#pragma unroll
        for (int q = 0; q < B; ++q)
#pragma unroll
          for (int r = 0; r < B; ++r) xb[r] = fma(-sD[B * b + r][B * c + q], xc[q], xb[r]);
      }
#pragma unroll
      for (int q = 0; q < B - 1; ++q)
#pragma unroll
        for (int r = q + 1; r < B; ++r) xb[r] = fma(-sD[B * b + r][B * b + q], xb[q], xb[r]);
#pragma unroll
      for (int r = 0; r < B; ++r) sX[B * b + r][lane] = xb[r];
    }
  } else {
    // row `lane`: x_k -= x_j D_jk for j < k, j increasing; x_k *= 1 / D_kk
#pragma unroll 1
    for (int b = 0; b < T / B; ++b) {
      S xb[B];
#pragma unroll
      for (int r = 0; r < B; ++r) xb[r] = sX[lane][B * b + r];
#pragma unroll 1
      for (int c = 0; c < b; ++c) {
        S xc[B];
#pragma unroll
        for (int q = 0; q < B; ++q) xc[q] = sX[lane][B * c + q];
#pragma unroll
        for (int q = 0; q < B; ++q)
#pragma unroll
          for (int r = 0; r < B; ++r) xb[r] = fma(-xc[q], sD[B * c + q][B * b + r], xb[r]);
      }
#pragma unroll
      for (int q = 0; q < B; ++q) {
        xb[q] *= rdiag[B * b + q];
#pragma unroll
        for (int r = q + 1; r < B; ++r) xb[r] = fma(-xb[q], sD[B * b + q][B * b + r], xb[r]);
      }
#pragma unroll
      for (int r = 0; r < B; ++r) sX[lane][B * b + r] = xb[r];
    }
  }
}

// Panel of step K, one wavefront per tile.  kind 0: U_KJ = L_KK^-1 P_K A_KJ
// (lane = column); kind 1: L_IK = A_IK U_KK^-1 (lane = row); kind 2: P_K
// applied to the L tiles already stored left of the diagonal in tile row K.
// The factored diagonal tile is staged in LDS and read with wave-uniform
// (broadcast) addresses.
template <class S>
__device__ __forceinline__ void panel_body(int64_t bid, int64_t K, int64_t nt, int tb, S *__restrict__ band,
                                              const int32_t *__restrict__ piv) {
  __shared__ __align__(16) S sD[T][T];
  __shared__ S sX[T][T + 1];
  __shared__ S rdiag[T];
  const int lane = threadIdx.x;
  const int kind = bid / tb, off = bid % tb;
  const int32_t *pv = piv + K * T;
  if (kind == 2) {
    const int64_t J = K - tb + off;
    if (J < 0) return;
    S *A = tile_ptr(band, tb, K, J);
    S x[T];
#pragma unroll
    for (int i = 0; i < T; ++i) x[i] = A[(int64_t)pv[i] * T + lane];
#pragma unroll
    for (int i = 0; i < T; ++i) A[i * T + lane] = x[i];
    return;
  }
  const int64_t IJ = K + 1 + off;
  if (IJ >= nt) return;
  const S *D = tile_ptr(band, tb, K, K);
  // full unrolling keeps all 64 row loads of a tile in flight (one wave per
  // tile: grouped loads would pay the memory latency once per group)
#pragma unroll
  for (int i = 0; i < T; ++i) sD[i][lane] = D[i * T + lane];
  S *A = kind == 0 ? tile_ptr(band, tb, K, IJ) : tile_ptr(band, tb, IJ, K);
#pragma unroll
  for (int i = 0; i < T; ++i) sX[i][lane] = A[(int64_t)(kind == 0 ? pv[i] : i) * T + lane];
  __syncthreads();
  if (kind == 1) {
    rdiag[lane] = S(1) / sD[lane][lane];
    __syncthreads();
  }
  trsm_wave<S>(kind, sD, sX, rdiag);
  __syncthreads();
#pragma unroll 8
  for (int i = 0; i < T; ++i) A[i * T + lane] = sX[i][lane];
}

template <class S>
__global__ __launch_bounds__(64) void k_panel(int64_t K, int64_t nt, int tb, S *__restrict__ band,
                                              const int32_t *__restrict__ piv) {
  panel_body<S>((int64_t)blockIdx.x, K, nt, tb, band, piv);
}

// C = A B for row-major 64x64 tiles in global memory, on the matrix cores
// (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32).  Both tiles are staged in
// LDS with 16-byte coalesced loads (rows padded to 65 entries, so fragment
// reads hit at most two lanes per bank); the 4 wavefronts of a 256-thread
// block take 16 rows each.  For each 16x16 output block and MFMA step kk,
// lane l (group g = l/16) feeds A[r0 + l%16][16g + kk] and
// B[16g + kk][c0 + l%16]: over the 16 steps the four groups cover k = 0..63.
// Result j = 0..3 of lane l lands at C[r0 + 4(l/16) + j][c0 + l%16] (FP32)
// or C[r0 + l/16 + 4j][c0 + l%16] (FP64), layouts checked on the device.
typedef double mm_d4 __attribute__((ext_vector_type(4)));
typedef float mm_f4 __attribute__((ext_vector_type(4)));
template <class S> struct Mfma;
template <> struct Mfma<double> {
  using V = mm_d4;
  static __device__ __forceinline__ V run(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma<float> {
  using V = mm_f4;
  static __device__ __forceinline__ V run(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// row / column of c[bj][j] in the 64x64 product
template <class S>
__device__ __forceinline__ int mm_row(int j) {
  const int q = (int)((threadIdx.x & 63) >> 4);
  return 16 * (int)(threadIdx.x >> 6) + (sizeof(S) == 8 ? q + 4 * j : 4 * q + j);
}
template <class S>
__device__ __forceinline__ int mm_col(int bj) { return 16 * bj + (int)(threadIdx.x & 15); }

constexpr int MM_LDS = 2 * T * (T + 1);  // entries of the LDS buffer of tile_mm

template <class S>
__device__ __forceinline__ void tile_mm(const S *__restrict__ a, const S *__restrict__ b,
                                        S c[4][4], S *lds) {
  using V = typename Mfma<S>::V;
  constexpr int P = T + 1;
  constexpr int W = 16 / sizeof(S);  // entries per 16-byte load
  S *sa = lds, *sb = lds + T * P;
  struct alignas(16) Vec { S v[W]; };
  Vec va[T * T / W / 256], vb[T * T / W / 256];
#pragma unroll
  for (int it = 0; it < T * T / W / 256; ++it) {
    const int e = (it * 256 + (int)threadIdx.x) * W;
    va[it] = *(const Vec *)(a + e);
    vb[it] = *(const Vec *)(b + e);
  }
#pragma unroll
  for (int it = 0; it < T * T / W / 256; ++it) {
    const int e = (it * 256 + (int)threadIdx.x) * W;
    const int r = e / T, col = e % T;
#pragma unroll
    for (int q = 0; q < W; ++q) {
      sa[r * P + col + q] = va[it].v[q];
      sb[r * P + col + q] = vb[it].v[q];
    }
  }
  __syncthreads();
  const int w = (int)(threadIdx.x >> 6), l = (int)(threadIdx.x & 63);
  const S *ar = sa + (16 * w + (l & 15)) * P + 16 * (l >> 4);
  const S *br = sb + (16 * (l >> 4)) * P + (l & 15);
  V acc[4];
#pragma unroll
  for (int bj = 0; bj < 4; ++bj) acc[bj] = V{0, 0, 0, 0};
#pragma unroll
  for (int kk = 0; kk < 16; ++kk) {
    const S av = ar[kk];
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) acc[bj] = Mfma<S>::run(av, br[kk * P + 16 * bj], acc[bj]);
  }
#pragma unroll
  for (int bj = 0; bj < 4; ++bj)
#pragma unroll
    for (int j = 0; j < 4; ++j) c[bj][j] = acc[bj][j];
}

// tile_mm with the A operands read straight into registers (lane l of
// wavefront w: row 16 w + l % 16, columns 16 (l / 16) .. + 15, contiguous in
// the row-major tile) and only B staged in LDS: MM1_LDS entries instead of
// MM_LDS, the same MFMA sequence and operands (bit-identical products).
constexpr int MM1_LDS = T * (T + 1) + 2 * T;  // B (or a staged tile) + two index arrays
template <class S>
__device__ __forceinline__ void tile_mm_ga(const S *__restrict__ a, const S *__restrict__ b,
                                           S c[4][4], S *lds, const int *brow = nullptr) {
  using V = typename Mfma<S>::V;
  constexpr int P = T + 1;
  constexpr int W = 16 / sizeof(S);  // entries per 16-byte load
  S *sb = lds;
  struct alignas(16) Vec { S v[W]; };
  const int w = (int)(threadIdx.x >> 6), l = (int)(threadIdx.x & 63);
  Vec ra[16 / W];
  const S *ag = a + (16 * w + (l & 15)) * T + 16 * (l >> 4);
#pragma unroll
  for (int q = 0; q < 16 / W; ++q) ra[q] = *(const Vec *)(ag + q * W);
  Vec vb[T * T / W / 256];
#pragma unroll
  for (int it = 0; it < T * T / W / 256; ++it) {
    const int e = (it * 256 + (int)threadIdx.x) * W;
    // brow: row r of the B operand is row brow[r] of b
    vb[it] = *(const Vec *)(brow ? b + brow[e / T] * T + e % T : b + e);
  }
#pragma unroll
  for (int it = 0; it < T * T / W / 256; ++it) {
    const int e = (it * 256 + (int)threadIdx.x) * W;
    const int r = e / T, col = e % T;
#pragma unroll
    for (int q = 0; q < W; ++q) sb[r * P + col + q] = vb[it].v[q];
  }
  __syncthreads();
  const S *br = sb + (16 * (l >> 4)) * P + (l & 15);
  V acc[4];
#pragma unroll
  for (int bj = 0; bj < 4; ++bj) acc[bj] = V{0, 0, 0, 0};
#pragma unroll
  for (int kk = 0; kk < 16; ++kk) {
    const S av = ra[kk / W].v[kk % W];
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) acc[bj] = Mfma<S>::run(av, br[kk * P + 16 * bj], acc[bj]);
  }
#pragma unroll
  for (int bj = 0; bj < 4; ++bj)
#pragma unroll
    for (int j = 0; j < 4; ++j) c[bj][j] = acc[bj][j];
}

// to_lds: the updated tile goes to lds as rows of T + 1 entries (for the
// diagonal LU that follows in the same block) instead of to the band.
// GA: tile_mm_ga (lds of MM1_LDS entries suffices).
template <class S, bool GA = false>
__device__ __forceinline__ void update_body(int64_t bid, int64_t K, int64_t nt, int tb, S *__restrict__ band,
                                            S *lds, bool to_lds = false) {
  const int64_t I = K + 1 + bid / tb, J = K + 1 + bid % tb;
  if (I >= nt || J >= nt) return;
  const S *Lt = tile_ptr(band, tb, I, K);
  const S *Ut = tile_ptr(band, tb, K, J);
  S *A = tile_ptr(band, tb, I, J);
  S old[4][4];  // the target tile is fetched together with the operands
#pragma unroll
  for (int bj = 0; bj < 4; ++bj)
#pragma unroll
    for (int j = 0; j < 4; ++j) old[bj][j] = A[mm_row<S>(j) * T + mm_col<S>(bj)];
  S c[4][4];
  if (GA) tile_mm_ga<S>(Lt, Ut, c, lds);
  else tile_mm<S>(Lt, Ut, c, lds);
  if (to_lds) {
    __syncthreads();  // every wavefront is done with the staged operands
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int j = 0; j < 4; ++j) lds[mm_row<S>(j) * (T + 1) + mm_col<S>(bj)] = old[bj][j] - c[bj][j];
    return;
  }
#pragma unroll
  for (int bj = 0; bj < 4; ++bj)
#pragma unroll
    for (int j = 0; j < 4; ++j) A[mm_row<S>(j) * T + mm_col<S>(bj)] = old[bj][j] - c[bj][j];
}

// accumulator row j of lane l in a 16 x 16 MFMA block (the FP64 and FP32
// instructions lay their results out differently)
template <class S>
__device__ __forceinline__ int acc_row(int j) {
  const int g = (int)((threadIdx.x & 63) >> 4);
  return sizeof(S) == 8 ? g + 4 * j : 4 * g + j;
}

// LU with partial pivoting of the 64x64 diagonal tile by a 256-thread block.
// The tile is in st (st[i][0..63] = row i, original order).  Rows are not
// moved: the candidate row with the largest |a_ij| (pivot_key) is the pivot
// of step j, and the row chosen at step j is stored as row j of the factored
// tile (P A = L U).  Panels of 16 columns: the first wavefront holds the
// panel's entries of every row in registers (lane = row) and runs its 16
// pivot steps (the pivot row's entries broadcast with readlane; rows no
// longer candidates take a zero multiplier); the panel's pivot rows then
// solve U12 = L11^-1 A12 (one trailing column per lane), and the four
// wavefronts subtract L21 U12 from the rows still unchosen on the matrix cores
// (wavefront w: rows 16w..16w+15; chosen rows enter with a zero multiplier
// and are not written).  Only the pivot steps stay on one wavefront.  The
// factored tile goes to A in pivot order and piv[K][j] records the original
// index of row j; prow / cand: 64 ints of LDS each (the row chosen at step j;
// whether row i is still a candidate).
// amax (used when PERT): static pivoting.  amax[0] holds max |K| as the bits
// of a non-negative double; a pivot of magnitude below thr = eps max |K| (eps
// of the storage precision: numerically zero) is replaced by +-thr (the
// factorization is then that of a nearby matrix, and the solve finishes with
// FGMRES) and counted in amax[1].
// 1 / x from the hardware reciprocal and two Newton steps (within about an
// ulp; the multipliers of the diagonal LU are scaled by it, as LAPACK's
// unblocked LU does; 0 gives inf as division would)
__device__ __forceinline__ double recip_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(fma(-x, y, 1.0), y, y);
  return fma(fma(-x, y, 1.0), y, y);
}
__device__ __forceinline__ float recip_nr(float x) {
  float y = __builtin_amdgcn_rcpf(x);
  return fmaf(fmaf(-x, y, 1.0f), y, y);
}

// Publication of one panel of a factored diagonal tile for the look-ahead
// panel solves (PUBW staging words): rows c0..c0+15 of the factors in pivot
// order (row j = original row prow[j], all 64 columns, final once its panel's
// pivot steps and U12 are done) and their pivots.  256 threads.
template <class S>
__device__ __forceinline__ void pub_stage(unsigned long long *__restrict__ pub, S (*st)[T + 1],
                                          const int *prow, int c0) {
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
#pragma unroll
  for (int it = 0; it < 4; ++it) {
    const int i = c0 + 4 * it + w;
    __hip_atomic_store(&pub[i * T + lane], pub_bits<S>(st[prow[i]][lane]), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  if (threadIdx.x < 16)
    __hip_atomic_store(&pub[T * T + c0 + threadIdx.x], (unsigned long long)prow[c0 + threadIdx.x],
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class S, bool PERT = false>
__device__ __forceinline__ void diag_core_mm(int64_t K, S *__restrict__ A, int32_t *__restrict__ piv,
                                             S (*st)[T + 1], int *prow, int *cand,
                                             unsigned long long *amax = nullptr,
                                             unsigned long long *pub = nullptr) {
  using V = typename Mfma<S>::V;
  const int t = (int)threadIdx.x, lane = t & 63, w = t >> 6;
  const int g = lane >> 4, cc = lane & 15;
  S thr = S(0);
  if (PERT) {
    const double rel = sizeof(S) == 8 ? 2.220446049250313e-16 : 1.1920928955078125e-7;
    thr = (S)(rel * __longlong_as_double((long long)__hip_atomic_load(
                           amax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
  }
  int npert = 0;
  int pos = -1;
#pragma unroll 1
  for (int pb = 0; pb < 4; ++pb) {
    const int c0 = pb * 16;
    if (w == 0) {
      S r[16];
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) r[jj] = st[lane][c0 + jj];
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) {
        const unsigned key = wave_max_u32(pivot_key(r[jj], pos < 0, lane));
        const int p = 63 - (int)(key & 63u);
        S pj = lane_bcast(r[jj], p);
        bool tiny = false;
        if (PERT) {  // wave-uniform
          tiny = fabs(pj) < thr;
          pj = tiny ? (pj < S(0) ? -thr : thr) : pj;
          npert += tiny ? 1 : 0;
        }
        if (lane == p) pos = c0 + jj;
        const S l = (pos < 0) ? r[jj] * recip_nr(pj) : S(0);
        if (pos < 0) r[jj] = l;
        if (PERT && tiny && lane == p) r[jj] = pj;
#pragma unroll
        for (int k = jj + 1; k < 16; ++k) r[k] = fma(-l, lane_bcast(r[k], p), r[k]);
      }
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) st[lane][c0 + jj] = r[jj];
      if (pos >= c0) prow[pos] = lane;
      cand[lane] = pos < 0 ? 1 : 0;
    }
    __syncthreads();
    if (pb == 3) {
      if (pub) pub_stage<S>(pub, st, prow, c0);
      break;
    }
    if (w == 0 && lane < 48 - c0) {  // U12 = L11^-1 A12, lane = trailing column
      const int c = c0 + 16 + lane;
      S u[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int ri = prow[c0 + i];
        S acc = st[ri][c];
#pragma unroll
        for (int ii = 0; ii < i; ++ii) acc = fma(-st[ri][c0 + ii], u[ii], acc);
        u[i] = acc;
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) st[prow[c0 + i]][c] = u[i];
    }
    __syncthreads();
    if (pub) pub_stage<S>(pub, st, prow, c0);  // this panel's pivot rows are final
    {  // A22 -= L21 U12 on the matrix cores, wavefront w: rows 16w .. 16w + 15
      const int row = 16 * w + cc;
      const bool cd = cand[row] != 0;
      S a[4];
      int br[4];
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int k = 4 * g + kk;
        a[kk] = cd ? -st[row][c0 + k] : S(0);
        br[kk] = prow[c0 + k];
      }
      bool wr[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) wr[j] = cand[16 * w + acc_row<S>(j)] != 0;
#pragma unroll 1
      for (int cb = 0; cb < 3 - pb; ++cb) {
        const int C0 = c0 + 16 + 16 * cb;
        V acc;
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = st[16 * w + acc_row<S>(j)][C0 + cc];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) acc = Mfma<S>::run(a[kk], st[br[kk]][C0 + cc], acc);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (wr[j]) st[16 * w + acc_row<S>(j)][C0 + cc] = acc[j];
      }
    }
    __syncthreads();
  }
  if (w == 0) {
    piv[K * T + pos] = lane;
    if (PERT && npert > 0 && lane == 0) atomicAdd(amax + 1, (unsigned long long)npert);
  }
#pragma unroll
  for (int it = 0; it < T / 4; ++it) {
    const int i = 4 * it + w;
    A[i * T + lane] = st[prow[i]][lane];
  }
}

// diag_core_mm on the tile staged in the first T (T + 1) entries of an LDS
// buffer of MM_LDS entries (its index arrays follow the tile)
template <class S, bool PERT>
__device__ __forceinline__ void diag_mm_lds(int64_t K, S *__restrict__ A, int32_t *__restrict__ piv,
                                            S *lds, unsigned long long *amax,
                                            unsigned long long *pub = nullptr) {
  S(*st)[T + 1] = reinterpret_cast<S(*)[T + 1]>(lds);
  int *prow = reinterpret_cast<int *>(lds + T * (T + 1));
  diag_core_mm<S, PERT>(K, A, piv, st, prow, prow + T, amax, pub);
}

// In-place inverses of the factors of a 64x64 diagonal tile factored by
// diag_core_mm (row i of P A = L U at st[prow[i]]: L unit lower below the
// diagonal, U upper), in 16 x 16 blocks.  The rows are first put in pivot
// order.  The diagonal blocks by column sweeps in place, one row per lane
// (wavefronts 0-1: L_bb^-1 of blocks 2w and 2w + 1, wavefronts 2-3: U_bb^-1),
// then three rounds on the matrix cores: round r forms block row r + 1 of
// X = L^-1, X_ij = -X_ii sum_{k=j}^{i-1} L_ik X_kj, and block row 2 - r of
// Y = U^-1, Y_ij = -Y_ii sum_{k=i+1}^{j} U_ik Y_kj, each from blocks that are
// final (every sum is formed before any job overwrites its block).  Li and Ui
// receive the inverses as full row-major tiles (zeros outside the triangles,
// ones on L^-1's diagonal).  256 threads; st is overwritten.
template <class S>
__device__ __forceinline__ int k_of(int kk) {  // the k a lane group feeds in MFMA step kk
  return acc_row<S>(kk);
}
template <class S>
__device__ __forceinline__ void tri_inv_lds(S (*st)[T + 1], const int *prow, S *__restrict__ Li,
                                            S *__restrict__ Ui) {
  using V = typename Mfma<S>::V;
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  const int cc = lane & 15;
  {  // rows into pivot order
    S v[T * T / 256];
#pragma unroll
    for (int it = 0; it < T * T / 256; ++it) {
      const int e = it * 256 + (int)threadIdx.x;
      v[it] = st[prow[e / T]][e % T];
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < T * T / 256; ++it) {
      const int e = it * 256 + (int)threadIdx.x;
      st[e / T][e % T] = v[it];
    }
    __syncthreads();
  }
  if (lane < 32) {
    // in place, row i = lane % 16 of block b per lane, columns in sequence
    // (a step's reads come before its writes: one wavefront per block pair)
    const int b0 = 16 * (2 * (w & 1) + (lane >> 4)), i = cc;
    S *ri = &st[b0 + i][b0];
    if (w < 2) {  // X = L_bb^-1, columns from the last: X[i][j] = -sum_{k=j+1}^{i} X[i][k] L[k][j]
#pragma unroll 1
      for (int j = 14; j >= 0; --j) {
        // every operand is loaded (in bounds), the terms outside j < k < i
        // enter as zeros: the loads issue together
        S a[14], b[14];
#pragma unroll
        for (int k = 1; k < 15; ++k) {
          a[k - 1] = ri[k];
          b[k - 1] = st[b0 + k][b0 + j];
        }
        const S lij = ri[j];
        S acc0 = i > j ? lij : S(0), acc1 = S(0);  // k = i: X[i][i] = 1
#pragma unroll
        for (int k = 1; k < 15; ++k) {
          const S t = (k > j && k < i) ? a[k - 1] : S(0);
          if (k & 1) acc1 = fma(t, b[k - 1], acc1);
          else acc0 = fma(t, b[k - 1], acc0);
        }
        if (i > j) ri[j] = -(acc0 + acc1);
      }
    } else {  // Y = U_bb^-1, columns from the first: Y[i][j] = -Y[j][j] sum_{k=i}^{j-1} Y[i][k] U[k][j]
#pragma unroll 1
      for (int j = 0; j < 16; ++j) {
        S a[15], b[15];
#pragma unroll
        for (int k = 0; k < 15; ++k) {
          a[k] = ri[k];
          b[k] = st[b0 + k][b0 + j];
        }
        const S yjj = S(1) / st[b0 + j][b0 + j];
        S acc0 = S(0), acc1 = S(0);
#pragma unroll
        for (int k = 0; k < 15; ++k) {
          const S t = (k >= i && k < j) ? a[k] : S(0);
          if (k & 1) acc1 = fma(t, b[k], acc1);
          else acc0 = fma(t, b[k], acc0);
        }
        if (i < j) ri[j] = -(acc0 + acc1) * yjj;
        else if (i == j) ri[j] = yjj;
      }
    }
  }
  __syncthreads();
  // entry (r, q) of block (bi, bj) of the tile
  auto at = [&](int bi, int bj, int r, int q) -> S { return st[16 * bi + r][16 * bj + q]; };
#pragma unroll 1
  for (int rd = 0; rd < 3; ++rd) {
    const int njob = 2 * (rd + 1);
    V tacc[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int q = w + 4 * h;
      tacc[h] = V{0, 0, 0, 0};
      if (q >= njob) continue;
      // the diagonal block of the sum (X_jj unit lower, Y_jj upper) is read
      // whole and masked
      if (q <= rd) {  // X row i = rd + 1, column j = q: sum_{k=j}^{i-1} L_ik X_kj
        const int i = rd + 1, j = q;
        for (int k = j; k < i; ++k)
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            const int km = k_of<S>(kk);
            const S a = at(i, k, cc, km), bv = at(k, j, km, cc);
            const S b = k > j || km > cc ? bv : (km == cc ? S(1) : S(0));
            tacc[h] = Mfma<S>::run(a, b, tacc[h]);
          }
      } else {  // Y row i = 2 - rd, column j = i + 1 + (q - rd - 1): sum_{k=i+1}^{j} U_ik Y_kj
        const int i = 2 - rd, j = i + 1 + (q - rd - 1);
        for (int k = i + 1; k <= j; ++k)
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            const int km = k_of<S>(kk);
            const S a = at(i, k, cc, km), bv = at(k, j, km, cc);
            const S b = k < j || km <= cc ? bv : S(0);
            tacc[h] = Mfma<S>::run(a, b, tacc[h]);
          }
      }
    }
    __syncthreads();  // every sum is formed before any block of this round is overwritten
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int q = w + 4 * h;
      if (q >= njob) continue;
      const bool isx = q <= rd;
      const int i = isx ? rd + 1 : 2 - rd, j = isx ? q : i + 1 + (q - rd - 1);
      V acc = V{0, 0, 0, 0};
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int km = k_of<S>(kk);
        // -X_ii (unit lower) or -Y_ii (upper); the sum's accumulator row
        // k_of(kk) of this lane's column is the B operand of step kk
        const S d = at(i, i, cc, km);
        const S a = isx ? (cc > km ? -d : (cc == km ? S(-1) : S(0))) : (cc <= km ? -d : S(0));
        acc = Mfma<S>::run(a, tacc[h][kk], acc);
      }
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) st[16 * i + acc_row<S>(jj)][16 * j + cc] = acc[jj];
    }
    __syncthreads();
  }
#pragma unroll 4
  for (int it = 0; it < T * T / 256; ++it) {
    const int e = it * 256 + (int)threadIdx.x, r = e / T, c = e % T;
    const S v = st[r][c];
    Li[e] = r > c ? v : (r == c ? S(1) : S(0));
    Ui[e] = r <= c ? v : S(0);
  }
}

// diagonal LU of tile (K, K) of a band by a 256-thread block
template <class S, bool PERT>
__device__ __forceinline__ void diag_block(int64_t K, int tb, S *__restrict__ band,
                                           int32_t *__restrict__ piv, unsigned long long *amax,
                                           S *lds) {
  S *A = tile_ptr(band, tb, K, K);
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
#pragma unroll
  for (int it = 0; it < T / 4; ++it) {
    const int i = 4 * it + w;
    lds[i * (T + 1) + lane] = A[i * T + lane];
  }
  __syncthreads();
  diag_mm_lds<S, PERT>(K, A, piv, lds, amax);
}

// the first diagonal LU of a chain
template <class S, bool PERT>
__global__ __launch_bounds__(256) void k_diag_mm(int64_t K, int tb, S *__restrict__ band,
                                                 int32_t *__restrict__ piv,
                                                 unsigned long long *__restrict__ amax) {
  __shared__ __align__(16) S lds[MM_LDS];
  diag_block<S, PERT>(K, tb, band, piv, amax, lds);
}

// Both chains of the two-ended factorization in one launch (blocks of chain 1
// follow those of chain 0).  A chain with fuse = 0 takes the plain trailing
// update (its next diagonal tile is not complete yet); nt = 0 leaves it out.
struct FacChain {
  void *band;
  int32_t *piv;
  int64_t K, nt;
  int32_t fuse;
  unsigned long long *pub;  // look-ahead staging, 2 PUBW words
  int64_t k0;               // the chain's first step (its row's L tiles permuted beforehand)
  // trailing tiles (I, J) with I, J >= s0 are not updated (a front's Schur
  // block, which receives all its steps' products at once afterwards)
  int64_t s0 = INT64_MAX;
};

template <class S>
__global__ __launch_bounds__(64) void k_panel2(int tb, FacChain c0, FacChain c1) {
  const int nb0 = 3 * tb;
  const bool second = (int)blockIdx.x >= nb0;
  const FacChain &ch = second ? c1 : c0;
  if (ch.nt == 0) return;
  panel_body<S>((int64_t)blockIdx.x - (second ? nb0 : 0), ch.K, ch.nt, tb, (S *)ch.band, ch.piv);
}

// ------------------------------------- factorization step with look-ahead
//
// One launch per tile step K of a chain does the trailing update of step K,
// the diagonal LU of step K + 1 and the panel of step K + 1, so the next
// step needs no separate panel launch:
//   block 0          : tile (K+1, K+1) updated and factored (diag_core_mm),
//                      its factors and pivots published to a staging buffer;
//   blocks 1 .. tb   : U panel tiles (K+1, J), J = K+2 .. K+1+tb: updated by
//                      step K when inside its window, then, once the factors of
//                      K + 1 are published, L^-1 P X (one wavefront);
//   tb+1 .. 2 tb     : L panel tiles (I, K+1) likewise, X U^-1;
//   2 tb+1 .. 3 tb   : P_(K+1) applied to the L tiles (K+1, J), J < K, and
//                      P_K to tile (K, K-1) (which the previous launch read
//                      as an operand, so could not permute);
//   the rest         : the other (tb - 1)^2 tiles of step K's window.
// Two staging buffers alternate with the step's parity; block 1 refills the
// one read in the previous launch with the sentinel.  Every waiting block
// waits only for block 0 of its chain, which is dispatched before it, so the
// waits end; each is bounded (on expiry the error count is raised and the
// block goes on with an invalid result).  A chain with fuse = 0 (its next
// diagonal tile is not complete) takes the plain trailing update.
constexpr int step_blocks(int tb) { return tb * tb + tb + 2; }

// waits until the staging words are published and copies them: the factored
// tile to sD (when set), the pivots to spiv.  One lane watches the last word
// first; then every thread issues all its loads at once and re-reads only
// words still holding the sentinel.
template <class S>
__device__ __forceinline__ void pub_wait(const unsigned long long *__restrict__ pub, S (*sD)[T],
                                         int *spiv, int *__restrict__ err, int *bad) {
  constexpr int NW = (PUBW + 255) / 256;
  if (threadIdx.x == 0) {
    *bad = 0;
    long spins = 0;
    while (__hip_atomic_load(&pub[PUBW - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
           SENTINEL) {
      __builtin_amdgcn_s_sleep(2);
      if (++spins > SPIN_LIMIT) {
        *bad = 1;
        atomicAdd(err, 1);
        break;
      }
    }
  }
  __syncthreads();
  if (*bad) return;
  const int i0 = (sD ? 0 : T * T) + (int)threadIdx.x;
  unsigned long long v[NW];
#pragma unroll
  for (int q = 0; q < NW; ++q) {
    const int idx = i0 + 256 * q;
    v[q] = idx < PUBW ? __hip_atomic_load(&pub[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                      : 0ull;
  }
#pragma unroll
  for (int q = 0; q < NW; ++q) {
    const int idx = i0 + 256 * q;
    if (idx >= PUBW) continue;
    long spins = 0;
    while (v[q] == SENTINEL) {
      __builtin_amdgcn_s_sleep(1);
      v[q] = __hip_atomic_load(&pub[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (++spins > SPIN_LIMIT) {
        atomicAdd(err, 1);
        break;
      }
    }
    if (idx < T * T)
      sD[idx / T][idx % T] = pub_val<S>(v[q]);
    else
      spiv[idx - T * T] = (int)(v[q] & 63);  // in range even for an expired wait
  }
  __syncthreads();
}

// P_I applied to the rows of the stored L tile (I, J) by a 256-thread block
// (all loads before the barrier, then the stores); pv: the 64 pivots of I
template <class S>
__device__ __forceinline__ void perm_rows(S *__restrict__ A, const int *pv) {
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  S v[T / 4];
#pragma unroll
  for (int it = 0; it < T / 4; ++it) v[it] = A[(int64_t)pv[4 * it + w] * T + lane];
  __syncthreads();
#pragma unroll
  for (int it = 0; it < T / 4; ++it) A[(4 * it + w) * T + lane] = v[it];
}

// P_I on tile (I, I - 1): the one L tile of row I that the look-ahead launch
// producing the factors of I cannot permute (it is an operand of that launch)
template <class S>
__global__ __launch_bounds__(256) void k_perm_tile(int tb, S *__restrict__ band,
                                                   const int32_t *__restrict__ piv, int64_t I) {
  __shared__ int pv[T];
  if (threadIdx.x < T) pv[threadIdx.x] = piv[I * T + threadIdx.x];
  __syncthreads();
  perm_rows<S>(tile_ptr(band, tb, I, I - 1), pv);
}

// Waits (wavefront 0 of a panel block) for stage st of the published factors
// (pivot rows 16 st .. 16 st + 15 and their pivots) and copies it: rows to
// sD (pivot order), pivots to spiv.  Bounded: an expired wait counts an error
// and returns false (the block's result is then invalid).
template <class S>
__device__ __forceinline__ bool stage_wait(const unsigned long long *__restrict__ pub, int st,
                                           S (*sD)[T], int *spiv, int *__restrict__ err) {
  const int lane = (int)(threadIdx.x & 63);
  long spins = 0;
  while (__hip_atomic_load(&pub[T * T + 16 * st + 15], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT) == SENTINEL) {
    __builtin_amdgcn_s_sleep(2);
    if (++spins > SPIN_LIMIT) {
      if (lane == 0) atomicAdd(err, 1);
      return false;
    }
  }
  unsigned long long v[16];
#pragma unroll
  for (int r = 0; r < 16; ++r)
    v[r] = __hip_atomic_load(&pub[(16 * st + r) * T + lane], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  unsigned long long pv = lane < 16 ? __hip_atomic_load(&pub[T * T + 16 * st + lane],
                                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                    : 0ull;
  bool ok = true;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    long sp = 0;
    while (v[r] == SENTINEL) {
      __builtin_amdgcn_s_sleep(1);
      v[r] = __hip_atomic_load(&pub[(16 * st + r) * T + lane], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      if (++sp > SPIN_LIMIT) {
        ok = false;
        break;
      }
    }
    sD[16 * st + r][lane] = pub_val<S>(v[r]);
  }
  if (lane < 16) {
    long sp = 0;
    while (pv == SENTINEL) {
      __builtin_amdgcn_s_sleep(1);
      pv = __hip_atomic_load(&pub[T * T + 16 * st + lane], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
      if (++sp > SPIN_LIMIT) {
        ok = false;
        break;
      }
    }
    spiv[16 * st + lane] = (int)(pv & 63);  // in range even for an expired wait
  }
  if (!__all(ok)) {
    if (lane == 0) atomicAdd(err, 1);
    return false;
  }
  __builtin_amdgcn_s_waitcnt(0);  // the LDS rows are read by other lanes next
  return true;
}

// The look-ahead panel solve, stage by stage as the diagonal LU publishes its
// panels, so all but the last stage overlap the LU.  Per stage (16 pivot
// rows / columns): the contributions of the earlier stages are subtracted on
// the matrix cores by the four wavefronts (left-looking), then wavefront 0
// solves the 16 x 16 diagonal block.  kind 0 (U tile, X <- L^-1 P X): X is
// the updated tile in its original row order and the result row i is stored
// in the slot of its source row spiv[i] (each source row is read once,
// before its slot is reused); wavefront w owns columns 16w..16w+15.  kind 1
// (L tile, X <- X U^-1): wavefront w owns rows 16w..16w+15.  256 threads;
// bad: an LDS int.  pub = null: sD and spiv already hold the whole factored
// tile and its pivots (no waits; bad is not read).
template <class S>
__device__ __forceinline__ void staged_solve(int kind, const unsigned long long *__restrict__ pub,
                                             S (*sD)[T], S (*sX)[T + 1], S *rdiag, int *spiv,
                                             int *__restrict__ err, int *bad) {
  using V = typename Mfma<S>::V;
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  const int g = lane >> 4, cc = lane & 15;
  for (int st = 0; st < 4; ++st) {
    const int s0 = 16 * st;
    if (pub && w == 0) {
      const bool ok = stage_wait<S>(pub, st, sD, spiv, err);
      if (lane == 0) *bad = ok ? 0 : 1;
    }
    __syncthreads();  // also orders the previous stage's diagonal solve before this stage
    if (pub && *bad) return;
    if (st > 0) {
      V acc;
      if (kind == 0) {  // rows s0.. (pivot order) -= L[s0.., 0..s0) Y[0..s0), columns 16w..
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = sX[spiv[s0 + acc_row<S>(j)]][16 * w + cc];
        for (int kb = 0; kb < st; ++kb)
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            const int k = 16 * kb + 4 * g + kk;
            acc = Mfma<S>::run(-sD[s0 + cc][k], sX[spiv[k]][16 * w + cc], acc);
          }
#pragma unroll
        for (int j = 0; j < 4; ++j) sX[spiv[s0 + acc_row<S>(j)]][16 * w + cc] = acc[j];
      } else {  // columns s0.. -= X[:, 0..s0) U[0..s0, s0..), rows 16w..
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = sX[16 * w + acc_row<S>(j)][s0 + cc];
        for (int kb = 0; kb < st; ++kb)
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            const int k = 16 * kb + 4 * g + kk;
            acc = Mfma<S>::run(-sX[16 * w + cc][k], sD[k][s0 + cc], acc);
          }
#pragma unroll
        for (int j = 0; j < 4; ++j) sX[16 * w + acc_row<S>(j)][s0 + cc] = acc[j];
      }
      __syncthreads();
    }
    if (w == 0) {
      S x[16];
      if (kind == 0) {  // unit lower 16 x 16 block, lane = column
#pragma unroll
        for (int r = 0; r < 16; ++r) x[r] = sX[spiv[s0 + r]][lane];
#pragma unroll
        for (int r = 1; r < 16; ++r)
#pragma unroll
          for (int q = 0; q < r; ++q) x[r] = fma(-sD[s0 + r][s0 + q], x[q], x[r]);
#pragma unroll
        for (int r = 0; r < 16; ++r) sX[spiv[s0 + r]][lane] = x[r];
      } else {  // upper 16 x 16 block from the right, lane = row
        if (lane < 16) rdiag[s0 + lane] = S(1) / sD[s0 + lane][s0 + lane];
        __builtin_amdgcn_s_waitcnt(0);
#pragma unroll
        for (int q = 0; q < 16; ++q) x[q] = sX[lane][s0 + q];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          x[q] *= rdiag[s0 + q];
#pragma unroll
          for (int r = q + 1; r < 16; ++r) x[r] = fma(-x[q], sD[s0 + q][s0 + r], x[r]);
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) sX[lane][s0 + q] = x[q];
      }
    }
  }
}

template <class S, bool PERT>
__device__ __forceinline__ void step_la_chain(int tb, const FacChain &ch, int64_t bid,
                                              unsigned long long *__restrict__ amax,
                                              int *__restrict__ err, S *lds, S *rdiag, int *spiv) {
  __shared__ int bad;
  S *band = (S *)ch.band;
  const int64_t K = ch.K, nt = ch.nt;
  if (!ch.fuse) {
    if (bid < (int64_t)tb * tb) {
      if (K + 1 + bid / tb >= ch.s0 && K + 1 + bid % tb >= ch.s0) return;
      update_body<S>(bid, K, nt, tb, band, lds);
    } else if (bid == (int64_t)tb * tb && K > ch.k0) {  // the deferred P_K of tile (K, K-1)
      if (threadIdx.x < T) spiv[threadIdx.x] = ch.piv[K * T + threadIdx.x];
      __syncthreads();
      perm_rows<S>(tile_ptr(band, tb, K, K - 1), spiv);
    }
    return;
  }
  unsigned long long *pub_new = ch.pub + ((K + 1) & 1) * PUBW;
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  if (bid == 0) {
    // the critical path of the launch: its wavefronts issue first on their CU
    __builtin_amdgcn_s_setprio(3);
    update_body<S>(0, K, nt, tb, band, lds, true);
    __syncthreads();
    diag_mm_lds<S, PERT>(K + 1, tile_ptr(band, tb, K + 1, K + 1), ch.piv, lds, amax, pub_new);
    return;
  }
  const int64_t r = bid - 1;
  if (r < 3 * (int64_t)tb) {
    if (r == 0) {  // refill the staging buffer read in the previous launch
      unsigned long long *pub_old = ch.pub + (K & 1) * PUBW;
      for (int idx = (int)threadIdx.x; idx < PUBW; idx += 256)
        __hip_atomic_store(&pub_old[idx], SENTINEL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const int kind = (int)(r / tb), off = (int)(r % tb);
    S(*sX)[T + 1] = reinterpret_cast<S(*)[T + 1]>(lds);
    S(*sD)[T] = reinterpret_cast<S(*)[T]>(lds + T * (T + 1));
    if (kind == 2) {
      if (off == tb - 1) {
        // tile (K+1, K) is step K's L operand of row K+1 in this very launch,
        // so its P_(K+1) waits for the next one; here: the P_K of (K, K-1)
        // deferred by the previous launch
        if (K <= ch.k0) return;
        if (threadIdx.x < T) spiv[threadIdx.x] = ch.piv[K * T + threadIdx.x];
        __syncthreads();
        perm_rows<S>(tile_ptr(band, tb, K, K - 1), spiv);
        return;
      }
      const int64_t J = K + 1 - tb + off;  // P_(K+1) on the L tile (K+1, J), J < K
      if (J < 0) return;
      pub_wait<S>(pub_new, nullptr, spiv, err, &bad);
      perm_rows<S>(tile_ptr(band, tb, K + 1, J), spiv);
      return;
    }
    const int64_t IJ = K + 2 + off;
    if (IJ >= nt) return;
    const int64_t I = kind == 0 ? K + 1 : IJ, J = kind == 0 ? IJ : K + 1;
    S *A = tile_ptr(band, tb, I, J);
    if (IJ <= K + tb) {  // inside step K's window: A_IJ - L_IK U_KJ, into LDS
      update_body<S>((I - K - 1) * tb + (J - K - 1), K, nt, tb, band, lds, true);
    } else {
#pragma unroll
      for (int it = 0; it < T / 4; ++it) sX[4 * it + w][lane] = A[(4 * it + w) * T + lane];
    }
    if (threadIdx.x < T) spiv[threadIdx.x] = (int)threadIdx.x;  // stays in range on an expired wait
    __syncthreads();
    staged_solve<S>(kind, pub_new, sD, sX, rdiag, spiv, err, &bad);
    __syncthreads();
#pragma unroll
    for (int it = 0; it < T / 4; ++it) {
      const int i = 4 * it + w;
      A[i * T + lane] = sX[kind == 0 ? spiv[i] : i][lane];
    }
    return;
  }
  const int64_t u = r - 3 * (int64_t)tb, tm = tb - 1;
  if (u >= tm * tm) return;
  const int64_t I = K + 2 + u / tm, J = K + 2 + u % tm;
  if (I >= ch.s0 && J >= ch.s0) return;
  update_body<S>((I - K - 1) * tb + (J - K - 1), K, nt, tb, band, lds);
}

// both chains of a two-ended launch: blocks of chain 1 follow chain 0's
template <class S, bool PERT>
__device__ __forceinline__ void step_la_body(int tb, const FacChain &c0, const FacChain &c1,
                                             unsigned long long *__restrict__ amax,
                                             int *__restrict__ err, S *lds, S *rdiag, int *spiv) {
  const int nbc = step_blocks(tb);
  const bool second = (int)blockIdx.x >= nbc;
  const FacChain &ch = second ? c1 : c0;
  if (ch.nt == 0) return;
  step_la_chain<S, PERT>(tb, ch, (int64_t)blockIdx.x - (second ? nbc : 0), amax, err, lds, rdiag,
                         spiv);
}

// span (the timed launch only, else null): the launch's own duration on the
// device's constant clock, as the earliest block start and latest block end
template <class S, bool PERT>
__global__ __launch_bounds__(256) void k_step_la(int tb, FacChain c0, FacChain c1,
                                                 unsigned long long *__restrict__ amax,
                                                 int *__restrict__ err,
                                                 unsigned long long *__restrict__ span) {
  __shared__ __align__(16) S lds[MM_LDS];
  __shared__ S rdiag[T];
  __shared__ int spiv[T];
  if (span && threadIdx.x == 0) atomicMin(&span[0], (unsigned long long)wall_clock64());
  step_la_body<S, PERT>(tb, c0, c1, amax, err, lds, rdiag, spiv);
  if (span) {
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(&span[1], (unsigned long long)wall_clock64());
  }
}

__global__ void k_span_init(unsigned long long *__restrict__ span) {
  if (threadIdx.x == 0) {
    span[0] = ~0ull;
    span[1] = 0ull;
  }
}

// ------------------------------------------------- post-factorization pass
//
// Inverts every factored diagonal tile (one wavefront per tile, lane = column
// of the inverse) and folds the inverses into the off-diagonal tiles,
// L~_IK = L_II^-1 L_IK and U~_IJ = U_II^-1 U_IJ, so that each sweep step is a
// single batch of tile mat-vecs: y_I = L_II^-1 P_I b_I - Σ L~_IK y_K and
// x_I = U_II^-1 y_I - Σ U~_IJ x_J.  Every tile is independent: fully parallel.

// Inverses of one factored diagonal tile by one wavefront: the tile staged in
// LDS, L^-1 = L^-1 I and U^-1 = I U^-1 by the panel's blocked substitutions
// (trsm_wave), so the factors are read from LDS with wave-uniform addresses.
template <class S>
__device__ __forceinline__ void tile_inv_body(int64_t bid, int tb, const S *__restrict__ band,
                                                 S *__restrict__ linv, S *__restrict__ uinv) {
  __shared__ __align__(16) S sD[T][T];
  __shared__ S sX[T][T + 1];
  __shared__ S rdiag[T];
  const int64_t I = bid;
  const int lane = threadIdx.x;
  const S *D = tile_ptr(band, tb, I, I);
#pragma unroll
  for (int i = 0; i < T; ++i) sD[i][lane] = D[i * T + lane];
#pragma unroll 8
  for (int i = 0; i < T; ++i) sX[i][lane] = i == lane ? S(1) : S(0);
  __syncthreads();
  trsm_wave<S>(0, sD, sX, rdiag);
  __syncthreads();
  S *Li = linv + I * (int64_t)T * T;
#pragma unroll 8
  for (int i = 0; i < T; ++i) Li[i * T + lane] = sX[i][lane];
  __syncthreads();
#pragma unroll 8
  for (int i = 0; i < T; ++i) sX[i][lane] = i == lane ? S(1) : S(0);
  rdiag[lane] = S(1) / sD[lane][lane];
  __syncthreads();
  trsm_wave<S>(1, sD, sX, rdiag);
  __syncthreads();
  S *Ui = uinv + I * (int64_t)T * T;
#pragma unroll 8
  for (int i = 0; i < T; ++i) Ui[i * T + lane] = sX[i][lane];
}

template <class S>
__global__ __launch_bounds__(64) void k_tile_inv(int tb, const S *__restrict__ band,
                                                 S *__restrict__ linv, S *__restrict__ uinv) {
  tile_inv_body<S>((int64_t)blockIdx.x, tb, band, linv, uinv);
}

// blockIdx.x = I (2 tb + 1 slots): slot < tb -> L tile (I, I - tb + slot),
// slot > tb -> U tile (I, I - tb + slot); the diagonal slot is skipped.
template <class S>
__device__ __forceinline__ void fold_body(int64_t bid, int64_t nt, int tb, S *__restrict__ band,
                                              const S *__restrict__ linv,
                                              const S *__restrict__ uinv, S *lds) {
  const int64_t I = bid / (2 * tb + 1);
  const int slot = bid % (2 * tb + 1);
  const int64_t J = I - tb + slot;
  if (slot == tb || J < 0 || J >= nt) return;
  const S *Inv = (slot < tb ? linv : uinv) + I * (int64_t)T * T;
  S *A = tile_ptr(band, tb, I, J);
  S c[4][4];
  tile_mm<S>(Inv, A, c, lds);  // reads all of A before its barrier
#pragma unroll
  for (int bj = 0; bj < 4; ++bj)
#pragma unroll
    for (int j = 0; j < 4; ++j) A[mm_row<S>(j) * T + mm_col<S>(bj)] = c[bj][j];
}

template <class S>
__global__ __launch_bounds__(256) void k_fold(int64_t nt, int tb, S *__restrict__ band,
                                              const S *__restrict__ linv,
                                              const S *__restrict__ uinv) {
  __shared__ __align__(16) S lds[MM_LDS];
  fold_body<S>((int64_t)blockIdx.x, nt, tb, band, linv, uinv, lds);
}

// ------------------------------------------------------------ triangular solves
//
// The band's tile rows form a dependency chain, so each sweep runs in one
// 1024-thread workgroup: no inter-workgroup synchronisation, the last tb
// solution tiles stay in an LDS ring, and each step streams the <= tb folded
// tiles of its row with 16 threads per tile row (4 columns each), all loads
// of a row in flight together.  The diagonal-block products (L_II^-1 P_I b_I
// before, U_II^-1 y_I between the sweeps) are independent per tile row and
// run fully parallel.

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

// Sum over the band tiles (I, K), K in [K0, K1), of tile_row4 against the ring;
// up to 12 (FP64) or 24 (FP32) tile loads per thread are issued before their
// results are used.
template <class S, int B = (sizeof(S) == 4 ? 24 : 10)>  // loads in flight within the VGPR budget
__device__ __forceinline__ double band_row_sum(const S *band, int tb, int64_t I, int64_t K0,
                                               int64_t K1, const double (*ring)[T],
                                               double acc = 0.0) {
  const int i = threadIdx.x >> 4, q = threadIdx.x & 15;
  for (int64_t Kb = K0; Kb < K1; Kb += B) {
    typename Vec4<S>::type m[B];
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (Kb + u < K1)
        m[u] = *reinterpret_cast<const typename Vec4<S>::type *>(
            tile_ptr(band, tb, I, Kb + u) + i * T + 4 * q);
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (Kb + u < K1) {
        const double *x = ring[(Kb + u) % RING] + 4 * q;
        acc = fma((double)m[u].x, x[0], acc);
        acc = fma((double)m[u].y, x[1], acc);
        acc = fma((double)m[u].z, x[2], acc);
        acc = fma((double)m[u].w, x[3], acc);
      }
  }
  return acc;
}

// The same sum over K = K1, K1 - 1, ..., K0 (decreasing), for the backward sweep.
template <class S, int B = (sizeof(S) == 4 ? 24 : 10)>
__device__ __forceinline__ double band_row_sum_desc(const S *band, int tb, int64_t I, int64_t K1,
                                                    int64_t K0, const double (*ring)[T],
                                                    double acc = 0.0) {
  const int i = threadIdx.x >> 4, q = threadIdx.x & 15;
  for (int64_t Kb = K1; Kb >= K0; Kb -= B) {
    typename Vec4<S>::type m[B];
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (Kb - u >= K0)
        m[u] = *reinterpret_cast<const typename Vec4<S>::type *>(
            tile_ptr(band, tb, I, Kb - u) + i * T + 4 * q);
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (Kb - u >= K0) {
        const double *x = ring[(Kb - u) % RING] + 4 * q;
        acc = fma((double)m[u].x, x[0], acc);
        acc = fma((double)m[u].y, x[1], acc);
        acc = fma((double)m[u].z, x[2], acc);
        acc = fma((double)m[u].w, x[3], acc);
      }
  }
  return acc;
}

template <class S>
__device__ __forceinline__ double tile_row_fma(const typename Vec4<S>::type &m, const double *x) {
  return fma((double)m.x, x[0], fma((double)m.y, x[1], fma((double)m.z, x[2], (double)m.w * x[3])));
}

// out_I = M_I v_I for every tile row (M = L^-1 with v = P b, or U^-1 with v = y)
// out_I = Inv_I sv (sv: the tile row's 64 values in LDS), four threads per
// output row, each summing 16 products in order before two shuffles
template <class S>
__device__ __forceinline__ void diag_mv(int64_t I, const S *__restrict__ inv, const double *sv,
                                        double *__restrict__ out) {
  const int i = threadIdx.x / 4, q = threadIdx.x % 4;
  const S *m = inv + I * (int64_t)T * T + i * T + q * 16;
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 16; ++k) s = fma((double)m[k], sv[q * 16 + k], s);
  s += __shfl_xor(s, 1, 64);
  s += __shfl_xor(s, 2, 64);
  if (q == 0) out[I * T + i] = s;
}

template <class S>
__device__ __forceinline__ void diag_apply_body(int64_t bid, const S *__restrict__ inv,
                                                    const int32_t *__restrict__ piv,
                                                    const double *__restrict__ v,
                                                    double *__restrict__ out,
                                                    bool atomic_in = false) {
  __shared__ double sv[T];
  const int64_t I = bid;
  if (threadIdx.x < T) {
    const int64_t k = I * T + (piv ? piv[I * T + threadIdx.x] : threadIdx.x);
    // the output of a multi-workgroup sweep is only ever accessed atomically
    sv[threadIdx.x] = atomic_in ? __longlong_as_double((long long)__hip_atomic_load(
                                      reinterpret_cast<const unsigned long long *>(v) + k,
                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                                : v[k];
  }
  __syncthreads();
  diag_mv<S>(I, inv, sv, out);
}

template <class S>
__global__ __launch_bounds__(256) void k_diag_apply(const S *__restrict__ inv,
                                                    const int32_t *__restrict__ piv,
                                                    const double *__restrict__ v,
                                                    double *__restrict__ out) {
  diag_apply_body<S>((int64_t)blockIdx.x, inv, piv, v, out);
}

template <class S>
__global__ __launch_bounds__(256) void k_diag_apply_atomic(const S *__restrict__ inv,
                                                           const double *__restrict__ v,
                                                           double *__restrict__ out) {
  diag_apply_body<S>((int64_t)blockIdx.x, inv, (const int32_t *)nullptr, v, out, true);
}

// y_I = c_I - Σ_{K<I} L~_IK y_K  (c precomputed into y).  Tiles K0 .. I - 2
// are summed in increasing order into one chain, the neighbour tile I - 1 on
// its own and added last: the order of the multi-workgroup sweep, so the
// result does not depend on the number of workgroups.
template <class S>
__device__ __forceinline__ void fwd_sweep_body(int64_t nt, int tb, const S *__restrict__ band,
                                               double *__restrict__ y) {
  __shared__ double ring[RING][T];
  const int i = threadIdx.x >> 4, q = threadIdx.x & 15;
  for (int64_t I = 0; I < nt; ++I) {
    const int64_t K0 = I - tb < 0 ? 0 : I - tb;
    double acc = band_row_sum(band, tb, I, K0, I - 1, ring);
    if (I >= 1)
      acc += tile_row_fma<S>(*reinterpret_cast<const typename Vec4<S>::type *>(
                                 tile_ptr(band, tb, I, I - 1) + i * T + 4 * q),
                             ring[(I - 1) % RING] + 4 * q);
    const double s = row16_sum(acc);
    if (q == 0) {
      const double u = y[I * T + i] - s;
      ring[I % RING][i] = u;
      y[I * T + i] = u;
    }
    __syncthreads();
  }
}

// x_I = d_I - Σ_{J>I} U~_IJ x_J  (d precomputed into x).  Tiles J1 .. I + 2
// are summed in decreasing order into one chain, the neighbour tile I + 1 on
// its own and added last (the order of the multi-workgroup sweep).
template <class S>
__device__ __forceinline__ void bwd_sweep_body(int64_t nt, int tb, const S *__restrict__ band,
                                               double *__restrict__ x) {
  __shared__ double ring[RING][T];
  const int i = threadIdx.x >> 4, q = threadIdx.x & 15;
  for (int64_t I = nt - 1; I >= 0; --I) {
    const int64_t J1 = I + tb > nt - 1 ? nt - 1 : I + tb;
    double acc = band_row_sum_desc<S, sizeof(S) == 4 ? 24 : 8>(band, tb, I, J1, I + 2, ring);
    if (I + 1 <= nt - 1)
      acc += tile_row_fma<S>(*reinterpret_cast<const typename Vec4<S>::type *>(
                                 tile_ptr(band, tb, I, I + 1) + i * T + 4 * q),
                             ring[(I + 1) % RING] + 4 * q);
    const double s = row16_sum(acc);
    if (q == 0) {
      const double u = x[I * T + i] - s;
      ring[I % RING][i] = u;
      x[I * T + i] = u;
    }
    __syncthreads();
  }
}

template <class S>
__global__ __launch_bounds__(SW) void k_fwd_sweep(int64_t nt, int tb, const S *__restrict__ band,
                                                  double *__restrict__ y) {
  fwd_sweep_body<S>(nt, tb, band, y);
}

// ---------------------------------------------------- multi-workgroup sweeps
//
// The same recurrences split over W workgroups: workgroup w solves tile rows
// w, w + W, ... (descending for the backward sweep).  Each solved row is
// published with relaxed agent-scope atomic stores into an output vector
// pre-filled with a sentinel, and consumers re-load a row until none of its
// 64 values is the sentinel: the value is its own ready flag, so no fences
// are needed.  A row's older tiles are loaded and summed while the row the
// chain waits for (the neighbour) is still being produced; only the last
// tile waits for it.  Every wait is bounded: on expiry the kernel counts an
// error and every workgroup leaves.


// The sweep outputs are written and read only with agent-scope atomics
// (never through a cached plain access), so no stale cache line can stand in
// for the sentinel or for a published row.
__global__ __launch_bounds__(256) void k_fill_sentinel(int64_t n, unsigned long long *__restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) __hip_atomic_store(&y[t], SENTINEL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// dst = src for a sweep output, read atomically, written as plain doubles
__global__ __launch_bounds__(256) void k_collect(int64_t n, const unsigned long long *__restrict__ src,
                                                 double *__restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n)
    dst[t] = __longlong_as_double((long long)__hip_atomic_load(&src[t], __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_AGENT));
}

// rows K of [k0, k1] into the LDS ring, one wave per row
__device__ __forceinline__ void stage_rows(const unsigned long long *__restrict__ y, int64_t k0,
                                           int64_t k1, double (*ring)[T], int *bad) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int64_t K = k0 + wave; K <= k1; K += SW / 64) {
    long spins = 0;
    unsigned long long v;
    for (;;) {
      v = __hip_atomic_load(&y[K * T + lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (__all(v != SENTINEL)) break;
      if (*bad) break;
      __builtin_amdgcn_s_sleep(1);
      if (++spins > SPIN_LIMIT) {
        *bad = 1;
        break;
      }
    }
    ring[K % RING][lane] = __longlong_as_double((long long)v);
  }
}

// y_I = c_I - Σ_{K<I} L~_IK y_K
template <class S>
__device__ __forceinline__ void fwd_sweep_mw_body(int w, int W, int64_t nt, int tb,
                                                  const S *__restrict__ band,
                                                  const double *__restrict__ c,
                                                  unsigned long long *__restrict__ y,
                                                  int *__restrict__ err) {
  __shared__ double ring[RING][T];
  __shared__ int bad;
  const int i = threadIdx.x >> 4, q = threadIdx.x & 15;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  int64_t have = -1;  // rows 0..have are in the ring
  for (int64_t I = w; I < nt; I += W) {
    const int64_t K0 = I - tb < 0 ? 0 : I - tb;
    typename Vec4<S>::type ml{};
    if (I >= 1)
      ml = *reinterpret_cast<const typename Vec4<S>::type *>(tile_ptr(band, tb, I, I - 1) + i * T +
                                                             4 * q);
    double acc = 0.0;
    // rows up to I - W are published (this workgroup's previous row and all
    // it staged before): their tiles are summed without waiting; rows
    // I - W + 1 .. I - 2 follow one at a time, each as it appears, in the
    // same increasing order, so the sum does not depend on W
    const int64_t Ka = I - W < I - 2 ? I - W : I - 2;
    if (Ka >= K0) {
      stage_rows(y, have + 1 > K0 ? have + 1 : K0, Ka, ring, &bad);
      if (Ka > have) have = Ka;
      __syncthreads();
      if (bad) break;
      acc = band_row_sum<S, sizeof(S) == 4 ? 12 : 8>(band, tb, I, K0, Ka + 1, ring);
    }
    for (int64_t r = Ka + 1 > K0 ? Ka + 1 : K0; r <= I - 2; ++r) {
      const typename Vec4<S>::type m =
          *reinterpret_cast<const typename Vec4<S>::type *>(tile_ptr(band, tb, I, r) + i * T + 4 * q);
      if (r > have) {
        stage_rows(y, r, r, ring, &bad);
        have = r;
      }
      __syncthreads();
      if (bad) break;
      const double *x = ring[r % RING] + 4 * q;
      acc = fma((double)m.x, x[0], acc);
      acc = fma((double)m.y, x[1], acc);
      acc = fma((double)m.z, x[2], acc);
      acc = fma((double)m.w, x[3], acc);
    }
    if (bad) break;
    if (I >= 1) {
      stage_rows(y, have + 1 > I - 1 ? have + 1 : I - 1, I - 1, ring, &bad);
      if (I - 1 > have) have = I - 1;
      __syncthreads();
      if (bad) break;
      acc += tile_row_fma<S>(ml, ring[(I - 1) % RING] + 4 * q);
    }
    const double s = row16_sum(acc);
    if (q == 0) {
      const double u = c[I * T + i] - s;
      __hip_atomic_store(&y[I * T + i], (unsigned long long)__double_as_longlong(u),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
  }
  if (bad && threadIdx.x == 0) atomicAdd(err, 1);
}

// x_I = d_I - Σ_{J>I} U~_IJ x_J
template <class S>
__device__ __forceinline__ void bwd_sweep_mw_body(int w, int W, int64_t nt, int tb,
                                                  const S *__restrict__ band,
                                                  const double *__restrict__ d,
                                                  unsigned long long *__restrict__ x,
                                                  int *__restrict__ err) {
  __shared__ double ring[RING][T];
  __shared__ int bad;
  const int i = threadIdx.x >> 4, q = threadIdx.x & 15;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  int64_t have = nt;  // rows have..nt-1 are in the ring
  for (int64_t I = nt - 1 - w; I >= 0; I -= W) {
    const int64_t J1 = I + tb > nt - 1 ? nt - 1 : I + tb;
    typename Vec4<S>::type ml{};
    if (I + 1 <= nt - 1)
      ml = *reinterpret_cast<const typename Vec4<S>::type *>(tile_ptr(band, tb, I, I + 1) + i * T +
                                                             4 * q);
    double acc = 0.0;
    // rows from I + W up are published (this workgroup's previous row and all
    // it staged before): their tiles are summed without waiting, highest
    // first; rows I + W - 1 .. I + 2 follow one at a time, each as it
    // appears, in the same decreasing order, so the sum does not depend on W
    const int64_t Ja = I + W > I + 2 ? I + W : I + 2;
    if (Ja <= J1) {
      const int64_t hi = have - 1 < J1 ? have - 1 : J1;
      stage_rows(x, Ja, hi, ring, &bad);
      if (Ja < have) have = Ja;
      __syncthreads();
      if (bad) break;
      acc = band_row_sum_desc<S, sizeof(S) == 4 ? 12 : 8>(band, tb, I, J1, Ja, ring);
    }
    for (int64_t r = (Ja - 1 < J1 ? Ja - 1 : J1); r >= I + 2; --r) {
      const typename Vec4<S>::type m =
          *reinterpret_cast<const typename Vec4<S>::type *>(tile_ptr(band, tb, I, r) + i * T + 4 * q);
      if (r < have) {
        stage_rows(x, r, r, ring, &bad);
        have = r;
      }
      __syncthreads();
      if (bad) break;
      const double *xr = ring[r % RING] + 4 * q;
      acc = fma((double)m.x, xr[0], acc);
      acc = fma((double)m.y, xr[1], acc);
      acc = fma((double)m.z, xr[2], acc);
      acc = fma((double)m.w, xr[3], acc);
    }
    if (bad) break;
    if (I + 1 <= nt - 1) {
      if (I + 1 < have) {
        stage_rows(x, I + 1, I + 1, ring, &bad);
        have = I + 1;
      }
      __syncthreads();
      if (bad) break;
      acc += tile_row_fma<S>(ml, ring[(I + 1) % RING] + 4 * q);
    }
    const double s = row16_sum(acc);
    if (q == 0) {
      const double u = d[I * T + i] - s;
      __hip_atomic_store(&x[I * T + i], (unsigned long long)__double_as_longlong(u),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
  }
  if (bad && threadIdx.x == 0) atomicAdd(err, 1);
}

template <class S>
__global__ __launch_bounds__(SW) void k_fwd_sweep_mw(int64_t nt, int tb, const S *__restrict__ band,
                                                     const double *__restrict__ c,
                                                     unsigned long long *__restrict__ y,
                                                     int *__restrict__ err) {
  fwd_sweep_mw_body<S>(blockIdx.x, gridDim.x, nt, tb, band, c, y, err);
}

template <class S>
__global__ __launch_bounds__(SW) void k_bwd_sweep_mw(int64_t nt, int tb, const S *__restrict__ band,
                                                     const double *__restrict__ d,
                                                     unsigned long long *__restrict__ x,
                                                     int *__restrict__ err) {
  bwd_sweep_mw_body<S>(blockIdx.x, gridDim.x, nt, tb, band, d, x, err);
}

template <class S>
__global__ __launch_bounds__(SW) void k_bwd_sweep(int64_t nt, int tb, const S *__restrict__ band,
                                                  double *__restrict__ x) {
  bwd_sweep_body<S>(nt, tb, band, x);
}

// The two chains of the two-ended solve (T in BT, B in BR) in one launch:
// chain = blockIdx.x / W.  Single workgroup (W = 1): in place on y.
// Multi-workgroup: c in, y out (sentinel-filled, atomic access only).
struct SweepChain {
  int64_t nt;
  const void *band;
  const double *c;
  double *y;
};

template <class S, int DIR>
__global__ __launch_bounds__(SW) void k_sweep2(int tb, SweepChain a0, SweepChain a1) {
  const SweepChain &a = blockIdx.x == 0 ? a0 : a1;
  if (DIR == 0)
    fwd_sweep_body<S>(a.nt, tb, (const S *)a.band, a.y);
  else
    bwd_sweep_body<S>(a.nt, tb, (const S *)a.band, a.y);
}

template <class S, int DIR>
__global__ __launch_bounds__(SW) void k_sweep2_mw(int tb, int W, SweepChain a0, SweepChain a1,
                                                  int *__restrict__ err) {
  const int chain = (int)blockIdx.x / W, w = (int)blockIdx.x % W;
  const SweepChain &a = chain == 0 ? a0 : a1;
  if (DIR == 0)
    fwd_sweep_mw_body<S>(w, W, a.nt, tb, (const S *)a.band, a.c,
                         reinterpret_cast<unsigned long long *>(a.y), err);
  else
    bwd_sweep_mw_body<S>(w, W, a.nt, tb, (const S *)a.band, a.c,
                         reinterpret_cast<unsigned long long *>(a.y), err);
}

// ------------------------------------------------------------ batched forms
//
// Patch p takes part in a factorization launch when on[p] && fac[p], and in
// a solve launch when its flag in the given mask is set.  Factorization steps locate their patch
// through the StepMap kernel argument; the other launches through segment
// maps.

__device__ __forceinline__ bool fac_on(const uint8_t *on, const uint8_t *fac, int p) {
  return on[p] && fac[p];
}

// storage precision a template instance serves; patches of the other
// precision skip the launch
template <class S>
constexpr int prec_of() {
  return sizeof(S) == 4 ? 32 : 64;
}

// 64 KiB chunks of a patch's band storage, 16 bytes per store
__global__ __launch_bounds__(256) void kb_zero_band(Seg sg, const PatchDev *__restrict__ pd,
                                                    const uint8_t *__restrict__ on,
                                                    const uint8_t *__restrict__ fac) {
  const int p = sg.map[blockIdx.x];
  if (!fac_on(on, fac, p)) return;
  const PatchDev &d = pd[p];
  const int64_t bytes = d.nt * (2 * d.tb + 1) * (int64_t)T * T * (d.prec / 8);
  const int64_t base = (int64_t)(blockIdx.x - sg.blk0[p]) * 65536;
  uint4 *dst = reinterpret_cast<uint4 *>(d.band);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int64_t off = base + k * 4096 + threadIdx.x * 16;
    if (off < bytes) dst[off / 16] = make_uint4(0u, 0u, 0u, 0u);
  }
}

template <class S>
__global__ __launch_bounds__(256) void kb_bsr_to_band(Seg sg, const PatchDev *__restrict__ pd,
                                                      const uint8_t *__restrict__ on,
                                                      const uint8_t *__restrict__ fac) {
  const int p = sg.map[blockIdx.x];
  if (!fac_on(on, fac, p) || pd[p].prec != prec_of<S>()) return;
  const PatchDev &d = pd[p];
  bsr_to_band_body<S>(blockIdx.x - sg.blk0[p], d.nnzb, d.ndf, d.tb, d.brow, d.colidx, d.bsr,
                      (S *)d.band);
}

// one block per patch: identity on the < 64 padding rows
template <class S>
__global__ __launch_bounds__(64) void kb_band_pad(const PatchDev *__restrict__ pd,
                                                  const uint8_t *__restrict__ on,
                                                  const uint8_t *__restrict__ fac) {
  const int p = blockIdx.x;
  if (!fac_on(on, fac, p) || pd[p].prec != prec_of<S>()) return;
  const PatchDev &d = pd[p];
  const int64_t R = d.nx + threadIdx.x;
  if (R < d.nt * T) tile_ptr((S *)d.band, d.tb, R / T, R / T)[(R % T) * T + (R % T)] = S(1);
}

// patch (through `order`) and block within the patch of a factorization step
// launch; per = blocks per patch as a function of its tb
template <int KIND>
__device__ __forceinline__ int step_locate(const StepMap &sm, const int32_t *__restrict__ order,
                                           int &local, int &tb) {
  const int b = blockIdx.x;
  int c = 0;
  while (c + 1 < sm.ncls && b >= sm.blk0[c + 1]) ++c;
  tb = sm.tb[c];
  const int per = KIND == 0 ? 1 : (KIND == 1 ? 3 * tb : tb * tb);
  const int rel = b - sm.blk0[c];
  local = rel % per;
  return order[sm.first[c] + rel / per];
}

template <class S>
__global__ __launch_bounds__(256) void kb_diag(StepMap sm, int64_t K, const int32_t *__restrict__ order,
                                               const PatchDev *__restrict__ pd,
                                               const uint8_t *__restrict__ on,
                                               const uint8_t *__restrict__ fac,
                                               const uint8_t *__restrict__ pert) {
  int local, tb;
  const int p = step_locate<0>(sm, order, local, tb);
  if (!fac_on(on, fac, p) || pd[p].prec != prec_of<S>()) return;
  __shared__ __align__(16) S lds[MM_LDS];
  if (pert[p])
    diag_block<S, true>(K, tb, (S *)pd[p].band, pd[p].piv, pd[p].amax, lds);
  else
    diag_block<S, false>(K, tb, (S *)pd[p].band, pd[p].piv, nullptr, lds);
}

// max |K| of every patch of precision S that factors with static pivoting,
// before its factorization: zeroed, then reduced over its block-sparse values
template <class S>
__global__ void kb_amax_zero(int P, const PatchDev *__restrict__ pd,
                             const uint8_t *__restrict__ on, const uint8_t *__restrict__ fac,
                             const uint8_t *__restrict__ pert) {
  const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (p < P && fac_on(on, fac, p) && pert[p] && pd[p].prec == prec_of<S>()) pd[p].amax[0] = 0;
}

template <class S>
__global__ __launch_bounds__(256) void kb_amax(Seg sg, const PatchDev *__restrict__ pd,
                                               const uint8_t *__restrict__ on,
                                               const uint8_t *__restrict__ fac,
                                               const uint8_t *__restrict__ pert) {
  const int p = sg.map[blockIdx.x];
  if (!fac_on(on, fac, p) || !pert[p] || pd[p].prec != prec_of<S>()) return;
  const PatchDev &d = pd[p];
  const int64_t n = d.nnzb * d.ndf * d.ndf;
  const int64_t t = (blockIdx.x - sg.blk0[p]) * 256 + threadIdx.x;
  double m = t < n ? fabs(d.bsr[t]) : 0.0;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0 && m > 0.0)
    atomicMax(d.amax, (unsigned long long)__double_as_longlong(m));
}

template <class S>
__global__ __launch_bounds__(64) void kb_panel(StepMap sm, int64_t K, const int32_t *__restrict__ order,
                                               const PatchDev *__restrict__ pd,
                                               const uint8_t *__restrict__ on,
                                               const uint8_t *__restrict__ fac) {
  int local, tb;
  const int p = step_locate<1>(sm, order, local, tb);
  if (!fac_on(on, fac, p) || pd[p].prec != prec_of<S>()) return;
  panel_body<S>(local, K, pd[p].nt, tb, (S *)pd[p].band, pd[p].piv);
}

template <class S>
__global__ __launch_bounds__(256) void kb_update(StepMap sm, int64_t K, const int32_t *__restrict__ order,
                                                 const PatchDev *__restrict__ pd,
                                                 const uint8_t *__restrict__ on,
                                                 const uint8_t *__restrict__ fac) {
  int local, tb;
  const int p = step_locate<2>(sm, order, local, tb);
  if (!fac_on(on, fac, p) || pd[p].prec != prec_of<S>()) return;
  __shared__ __align__(16) S lds[MM_LDS];
  update_body<S>(local, K, pd[p].nt, tb, (S *)pd[p].band, lds);
}

template <class S>
__global__ __launch_bounds__(64) void kb_tile_inv(Seg sg, const PatchDev *__restrict__ pd,
                                                  const uint8_t *__restrict__ on,
                                                  const uint8_t *__restrict__ fac) {
  const int p = sg.map[blockIdx.x];
  if (!fac_on(on, fac, p) || pd[p].prec != prec_of<S>()) return;
  const PatchDev &d = pd[p];
  tile_inv_body<S>(blockIdx.x - sg.blk0[p], d.tb, (const S *)d.band, (S *)d.linv, (S *)d.uinv);
}

template <class S>
__global__ __launch_bounds__(256) void kb_fold(Seg sg, const PatchDev *__restrict__ pd,
                                               const uint8_t *__restrict__ on,
                                               const uint8_t *__restrict__ fac) {
  const int p = sg.map[blockIdx.x];
  if (!fac_on(on, fac, p) || pd[p].prec != prec_of<S>()) return;
  const PatchDev &d = pd[p];
  __shared__ __align__(16) S lds[MM_LDS];
  fold_body<S>(blockIdx.x - sg.blk0[p], d.nt, d.tb, (S *)d.band, (const S *)d.linv,
               (const S *)d.uinv, lds);
}

// which = 0: t2 = L^-1 P w; which = 1: t = U^-1 t2 (per tile row)
template <class S>
__global__ __launch_bounds__(256) void kb_diag_apply(Seg sg, const PatchDev *__restrict__ pd,
                                                     const uint8_t *__restrict__ mask, int which) {
  const int p = sg.map[blockIdx.x];
  if (!mask[p] || pd[p].prec != prec_of<S>()) return;
  const PatchDev &d = pd[p];
  if (which == 0)
    diag_apply_body<S>(blockIdx.x - sg.blk0[p], (const S *)d.linv, d.piv, d.w, d.t2);
  else if (d.sweep_w > 1)  // forward result in s1 (atomic); the backward sweep reads t2
    diag_apply_body<S>(blockIdx.x - sg.blk0[p], (const S *)d.uinv, (const int32_t *)nullptr, d.s1,
                       d.t2, true);
  else
    diag_apply_body<S>(blockIdx.x - sg.blk0[p], (const S *)d.uinv, (const int32_t *)nullptr, d.t2,
                       d.t);
}

template <class S>
__global__ __launch_bounds__(SW) void kb_fwd_sweep(const PatchDev *__restrict__ pd,
                                                   const uint8_t *__restrict__ mask) {
  const int p = blockIdx.x;
  if (!mask[p] || pd[p].prec != prec_of<S>() || pd[p].sweep_w > 1) return;
  fwd_sweep_body<S>(pd[p].nt, pd[p].tb, (const S *)pd[p].band, pd[p].t2);
}

template <class S>
__global__ __launch_bounds__(SW) void kb_bwd_sweep(const PatchDev *__restrict__ pd,
                                                   const uint8_t *__restrict__ mask) {
  const int p = blockIdx.x;
  if (!mask[p] || pd[p].prec != prec_of<S>() || pd[p].sweep_w > 1) return;
  bwd_sweep_body<S>(pd[p].nt, pd[p].tb, (const S *)pd[p].band, pd[p].t);
}

// Patches with sweep_w > 1 sweep over sweep_w workgroups (segment `sg`);
// their outputs are sentinel-filled first.  which = 0: forward sweep of t2
// into s1; which = 1: backward sweep of t2 into s2 (then collected into t).
__global__ __launch_bounds__(256) void kb_fill_sentinel(Seg sg, const PatchDev *__restrict__ pd,
                                                        const uint8_t *__restrict__ mask, int which,
                                                        int prec) {
  const int p = sg.map[blockIdx.x];
  if (!mask[p] || pd[p].sweep_w <= 1 || pd[p].prec != prec) return;
  const PatchDev &d = pd[p];
  const int64_t t = (int64_t)(blockIdx.x - sg.blk0[p]) * 256 + threadIdx.x;
  if (t < d.nt * T)
    __hip_atomic_store(reinterpret_cast<unsigned long long *>(which == 0 ? d.s1 : d.s2) + t,
                       SENTINEL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// t = backward-sweep output (atomic reads) for multi-workgroup patches
__global__ __launch_bounds__(256) void kb_collect(Seg sg, const PatchDev *__restrict__ pd,
                                                  const uint8_t *__restrict__ mask, int prec) {
  const int p = sg.map[blockIdx.x];
  if (!mask[p] || pd[p].sweep_w <= 1 || pd[p].prec != prec) return;
  const PatchDev &d = pd[p];
  const int64_t t = (int64_t)(blockIdx.x - sg.blk0[p]) * 256 + threadIdx.x;
  if (t < d.nt * T)
    d.t[t] = __longlong_as_double((long long)__hip_atomic_load(
        reinterpret_cast<const unsigned long long *>(d.s2) + t, __ATOMIC_RELAXED,
        __HIP_MEMORY_SCOPE_AGENT));
}

template <class S>
__global__ __launch_bounds__(SW) void kb_sweep_mw(Seg sg, const PatchDev *__restrict__ pd,
                                                  const uint8_t *__restrict__ mask, int which) {
  const int p = sg.map[blockIdx.x];
  if (!mask[p] || pd[p].prec != prec_of<S>()) return;
  const PatchDev &d = pd[p];
  const int w = (int)(blockIdx.x - sg.blk0[p]);
  if (which == 0)
    fwd_sweep_mw_body<S>(w, d.sweep_w, d.nt, d.tb, (const S *)d.band, d.t2,
                         reinterpret_cast<unsigned long long *>(d.s1), d.err);
  else
    bwd_sweep_mw_body<S>(w, d.sweep_w, d.nt, d.tb, (const S *)d.band, d.t2,
                         reinterpret_cast<unsigned long long *>(d.s2), d.err);
}

// ------------------------------------------------------------------ host side

static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

template <class S>
static void fill_t(Ctx &c, S *band) {
  const int64_t nband = c.band_rows() * (2 * c.tb + 1) * (int64_t)T * T;
  // grid-stride over a capped grid: a dispatch counts its work-items in 32
  // bits, and a large patch's band (256x256 ALEVB: 4.5e9 entries) exceeds that
  const unsigned zb = (unsigned)std::min<int64_t>(nblk(nband), 1 << 20);
  hipLaunchKernelGGL(k_zero_t<S>, dim3(zb), dim3(256), 0, c.stream, nband, band);
  const int64_t nb = c.nnzb * c.ndf * c.ndf;
  const int64_t npad = c.nt * T;
  if (c.tw_a > 0) {
    const int64_t rows_t = c.tw_a + c.tb;  // BT: T and the separator
    S *br = band + rows_t * (2 * c.tb + 1) * (int64_t)T * T;
    hipLaunchKernelGGL(k_bsr_to_band_tw<S>, dim3(nblk(nb)), dim3(256), 0, c.stream, c.nnzb, c.ndf,
                       c.tb, rows_t * T, npad, c.d_brow, c.d_colidx, c.d_bsr, band, br);
    if (npad > c.nx)
      hipLaunchKernelGGL(k_band_pad_tw<S>, dim3((unsigned)((npad - c.nx + 63) / 64)), dim3(64), 0,
                         c.stream, c.nx, npad, c.tb, br);
    MAFEM_HIP(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(k_bsr_to_band<S>, dim3(nblk(nb)), dim3(256), 0, c.stream, c.nnzb, c.ndf, c.tb,
                     c.d_brow, c.d_colidx, c.d_bsr, band);
  if (npad > c.nx)
    hipLaunchKernelGGL(k_band_pad<S>, dim3((unsigned)((npad - c.nx + 63) / 64)), dim3(64), 0,
                       c.stream, c.nx, npad, c.tb, band);
  MAFEM_HIP(hipGetLastError());
}

// Steps k0 .. k1 - 1 of the band LU of a band with nt tile rows on stream s:
// the diagonal LU of step k0 first (diag_first), the panel of step k0, then
// one look-ahead launch per step (k_step_la: the trailing update, the next
// diagonal LU and the next panel), except after the last step when
// !fuse_last (its next diagonal tile is not complete yet: plain update).
// pub: the chain's 2 PUBW staging words.  When `timed`, one launch in the
// middle records its duration in c.d_span.
template <class S, bool PERT>
static void factor_chain(Ctx &c, hipStream_t s, S *band, int32_t *piv, unsigned long long *pub,
                         int64_t nt, int64_t k0, int64_t k1, bool diag_first, bool fuse_last,
                         bool timed) {
  if (diag_first)
    hipLaunchKernelGGL((k_diag_mm<S, PERT>), dim3(1), dim3(256), 0, s, k0, c.tb, band, piv, c.d_amax);
  hipLaunchKernelGGL(k_fill_sentinel, dim3((2 * PUBW + 255) / 256), dim3(256), 0, s,
                     (int64_t)(2 * PUBW), pub);
  hipLaunchKernelGGL(k_panel<S>, dim3(3 * c.tb), dim3(64), 0, s, k0, nt, c.tb, band, piv);
  const int64_t timed_K = timed ? (k0 + k1) / 2 : -1;
  int64_t last_lu = k0;  // the last diagonal factored by a look-ahead launch
  for (int64_t K = k0; K < k1; ++K) {
    if (K + 1 >= nt) break;  // no trailing tiles after the last step
    const FacChain ch{band, piv, K, nt, (K + 1 < k1 || fuse_last) ? 1 : 0, pub, k0};
    const FacChain none{band, piv, K, 0, 0, pub, k0};
    last_lu = ch.fuse ? K + 1 : k0;
    const bool tk = (K == timed_K);
    if (tk) hipLaunchKernelGGL(k_span_init, dim3(1), dim3(64), 0, s, c.d_span);
    hipLaunchKernelGGL((k_step_la<S, PERT>), dim3(step_blocks(c.tb)), dim3(256), 0, s, c.tb, ch,
                       none, c.d_amax, c.d_sweep_err, tk ? c.d_span : nullptr);
    if (tk) c.step_timed = true;
  }
  if (last_lu > k0)  // no later launch applies its P to tile (last_lu, last_lu - 1)
    hipLaunchKernelGGL(k_perm_tile<S>, dim3(1), dim3(256), 0, s, c.tb, band, piv, last_lu);
}

template <class S, bool PERT>
static void factor_run(Ctx &c, S *band, S *li, S *ui) {
  if (PERT) {  // max |K| of this tangent for the static-pivoting threshold
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c.stream, c.d_amax);
    const int64_t nb = c.nnzb * c.ndf * c.ndf;
    hipLaunchKernelGGL(k_absmax<double>, dim3(1024), dim3(256), 0, c.stream, nb,
                       (const double *)c.d_bsr, c.d_amax);
  }
  if (c.tw_a == 0) {
    factor_chain<S, PERT>(c, c.stream, band, c.d_piv, c.d_pub, c.nt, 0, c.nt, true, true, true);
    hipLaunchKernelGGL(k_tile_inv<S>, dim3((unsigned)c.nt), dim3(64), 0, c.stream, c.tb, band, li, ui);
    hipLaunchKernelGGL(k_fold<S>, dim3((unsigned)(c.nt * (2 * c.tb + 1))), dim3(256), 0, c.stream,
                       c.nt, c.tb, band, li, ui);
    MAFEM_HIP(hipGetLastError());
    return;
  }
  // two-ended: step k of T (downward) and of B (upward, reversed band) in the
  // same launches, then the separator, which has both Schur complements added
  const int64_t a = c.tw_a, nb = c.tw_nb, w = c.tb, rt = a + w, rb = nb + w;
  const int64_t tsz = (int64_t)T * T;
  S *br = band + rt * (2 * c.tb + 1) * tsz;
  int32_t *pr = c.d_piv + rt * T;
  hipLaunchKernelGGL((k_diag_mm<S, PERT>), dim3(1), dim3(256), 0, c.stream, (int64_t)0, c.tb, band,
                     c.d_piv, c.d_amax);
  hipLaunchKernelGGL((k_diag_mm<S, PERT>), dim3(1), dim3(256), 0, c.stream, (int64_t)0, c.tb, br, pr,
                     c.d_amax);
  unsigned long long *pt = c.d_pub, *pb = c.d_pub + 2 * PUBW;
  hipLaunchKernelGGL(k_fill_sentinel, dim3((4 * PUBW + 255) / 256), dim3(256), 0, c.stream,
                     (int64_t)(4 * PUBW), c.d_pub);
  const int64_t steps = std::max(a, nb), timed_k = std::min(a, nb) / 2;
  {
    const FacChain ct{band, c.d_piv, 0, a > 0 ? rt : 0, 0, pt, 0};
    const FacChain cb{br, pr, 0, nb > 0 ? rb : 0, 0, pb, 0};
    hipLaunchKernelGGL(k_panel2<S>, dim3(6 * c.tb), dim3(64), 0, c.stream, c.tb, ct, cb);
  }
  for (int64_t k = 0; k < steps; ++k) {
    // a chain past its last step drops out (nt = 0); each chain's last step
    // leaves its separator diagonal tile unfactored (fuse = 0)
    // (each chain's last step is a plain update, which also applies the
    // deferred P of its last factored tile row)
    const FacChain ct{band, c.d_piv, k, k < a ? rt : 0, k + 1 < a ? 1 : 0, pt, 0};
    const FacChain cb{br, pr, k, k < nb ? rb : 0, k + 1 < nb ? 1 : 0, pb, 0};
    const bool tk = (k == timed_k);
    if (tk) hipLaunchKernelGGL(k_span_init, dim3(1), dim3(64), 0, c.stream, c.d_span);
    hipLaunchKernelGGL((k_step_la<S, PERT>), dim3(2 * step_blocks(c.tb)), dim3(256), 0, c.stream,
                       c.tb, ct, cb, c.d_amax, c.d_sweep_err, tk ? c.d_span : nullptr);
    if (tk) c.step_timed = true;
  }
  hipLaunchKernelGGL(k_sep_add<S>, dim3((unsigned)(w * w)), dim3(256), 0, c.stream, a, (int)w, c.tb,
                     c.nt, band, (const S *)br);
  factor_chain<S, PERT>(c, c.stream, band, c.d_piv, pt, rt, a, rt, true, true, false);
  // inverses of the diagonal tiles; folds: every tile row of BT, B's rows of
  // BR (its separator rows keep their raw L tiles for the solve)
  hipLaunchKernelGGL(k_tile_inv<S>, dim3((unsigned)rt), dim3(64), 0, c.stream, c.tb, band, li, ui);
  hipLaunchKernelGGL(k_tile_inv<S>, dim3((unsigned)nb), dim3(64), 0, c.stream, c.tb, (const S *)br,
                     li + rt * tsz, ui + rt * tsz);
  hipLaunchKernelGGL(k_fold<S>, dim3((unsigned)(rt * (2 * c.tb + 1))), dim3(256), 0, c.stream, rt,
                     c.tb, band, li, ui);
  hipLaunchKernelGGL(k_fold<S>, dim3((unsigned)(nb * (2 * c.tb + 1))), dim3(256), 0, c.stream, rb,
                     c.tb, br, li + rt * tsz, ui + rt * tsz);
  MAFEM_HIP(hipGetLastError());
}

// Two-ended solve (Ctx::tw_a > 0).  Forward: c = L_II^-1 P b on the rows of T
// and (reversed) of B, both forward sweeps at once; the separator's
// right-hand side loses B's part (BR's raw separator L tiles against B's
// solution, reversed) before its own L_II^-1 P, then T's part (BT's folded
// tiles) and its own sweep.  Backward: U_II^-1 everywhere, the separator's
// backward sweep, its terms in the last tb rows of T and of B, then both
// backward sweeps at once.
template <class S>
static void solve_tw(Ctx &c, const S *band, const S *li, const S *ui, const double *d_rhs,
                     double *d_x) {
  const int64_t a = c.tw_a, nb = c.tw_nb, w = c.tb, rt = a + w, tsz = (int64_t)T * T;
  const int64_t N = c.nt * T, ns = rt * T;
  const int tb = c.tb;
  const S *bt = band, *br = band + rt * (2 * tb + 1) * tsz;
  const S *sep = bt + a * (2 * tb + 1) * tsz;  // BT from the separator's first row
  const S *liR = li + rt * tsz, *uiR = ui + rt * tsz;
  const int32_t *pivR = c.d_piv + rt * T;
  const hipStream_t s = c.stream;
  const bool mw = c.sweep_groups > 1;
  const int W = c.sweep_groups;
  hipLaunchKernelGGL(k_rev_copy<false>, dim3(nblk(nb * T)), dim3(256), 0, s, (int64_t)0, nb * T, N,
                     d_rhs, c.d_tw_br);
  hipLaunchKernelGGL(k_diag_apply<S>, dim3((unsigned)a), dim3(256), 0, s, li, c.d_piv, d_rhs, c.d_t2);
  hipLaunchKernelGGL(k_diag_apply<S>, dim3((unsigned)nb), dim3(256), 0, s, liR, pivR,
                     (const double *)c.d_tw_br, c.d_tw_cr);
  if (mw) {
    hipLaunchKernelGGL(k_fill_sentinel, dim3(nblk(a * T)), dim3(256), 0, s, a * T,
                       (unsigned long long *)c.d_s1);
    hipLaunchKernelGGL(k_fill_sentinel, dim3(nblk(nb * T)), dim3(256), 0, s, nb * T,
                       (unsigned long long *)c.d_tw_s1);
    hipLaunchKernelGGL((k_sweep2_mw<S, 0>), dim3(2 * W), dim3(SW), 0, s, tb, W,
                       SweepChain{a, bt, c.d_t2, c.d_s1}, SweepChain{nb, br, c.d_tw_cr, c.d_tw_s1},
                       c.d_sweep_err);
  } else {
    hipLaunchKernelGGL((k_sweep2<S, 0>), dim3(2), dim3(SW), 0, s, tb,
                       SweepChain{a, bt, nullptr, c.d_t2}, SweepChain{nb, br, nullptr, c.d_tw_cr});
  }
  const double *yT = mw ? c.d_s1 : c.d_t2, *yB = mw ? c.d_tw_s1 : c.d_tw_cr;
  MAFEM_HIP(hipMemcpyAsync(c.d_tw_z + a * T, d_rhs + a * T, sizeof(double) * w * T,
                           hipMemcpyDeviceToDevice, s));
  if (mw)
    hipLaunchKernelGGL((k_rows_sub<S, true>), dim3((unsigned)w), dim3(256), 0, s, tb, br, nb,
                       (int64_t)0, nb, yB, (int64_t)0, c.d_tw_z, N);
  else
    hipLaunchKernelGGL((k_rows_sub<S, false>), dim3((unsigned)w), dim3(256), 0, s, tb, br, nb,
                       (int64_t)0, nb, yB, (int64_t)0, c.d_tw_z, N);
  hipLaunchKernelGGL(k_diag_apply<S>, dim3((unsigned)w), dim3(256), 0, s, li + a * tsz,
                     c.d_piv + a * T, (const double *)c.d_tw_z + a * T, c.d_t2 + a * T);
  if (mw)
    hipLaunchKernelGGL((k_rows_sub<S, true>), dim3((unsigned)w), dim3(256), 0, s, tb, bt, a,
                       (int64_t)0, a, yT, (int64_t)0, c.d_t2, (int64_t)0);
  else
    hipLaunchKernelGGL((k_rows_sub<S, false>), dim3((unsigned)w), dim3(256), 0, s, tb, bt, a,
                       (int64_t)0, a, yT, (int64_t)0, c.d_t2, (int64_t)0);
  double *ys = c.d_t2 + a * T;  // the separator's rows, swept in place
  if (mw) {  // over several workgroups too (rows a.. of s1 / s2 are free here)
    unsigned long long *o = (unsigned long long *)c.d_s1 + a * T;
    hipLaunchKernelGGL(k_fill_sentinel, dim3(nblk(w * T)), dim3(256), 0, s, w * T, o);
    hipLaunchKernelGGL(k_fwd_sweep_mw<S>, dim3((unsigned)std::min<int64_t>(W, w)), dim3(SW), 0, s,
                       w, tb, sep, (const double *)ys, o, c.d_sweep_err);
    hipLaunchKernelGGL(k_collect, dim3(nblk(w * T)), dim3(256), 0, s, w * T,
                       (const unsigned long long *)o, ys);
  } else {
    hipLaunchKernelGGL(k_fwd_sweep<S>, dim3(1), dim3(SW), 0, s, w, tb, sep, ys);
  }
  // d = U_II^-1 y (in place per tile row where y is plain)
  if (mw) {
    hipLaunchKernelGGL(k_diag_apply_atomic<S>, dim3((unsigned)a), dim3(256), 0, s, ui, yT, c.d_t2);
    hipLaunchKernelGGL(k_diag_apply_atomic<S>, dim3((unsigned)nb), dim3(256), 0, s, uiR, yB,
                       c.d_tw_dr);
  } else {
    hipLaunchKernelGGL(k_diag_apply<S>, dim3((unsigned)a), dim3(256), 0, s, ui,
                       (const int32_t *)nullptr, yT, c.d_t2);
    hipLaunchKernelGGL(k_diag_apply<S>, dim3((unsigned)nb), dim3(256), 0, s, uiR,
                       (const int32_t *)nullptr, yB, c.d_tw_dr);
  }
  hipLaunchKernelGGL(k_diag_apply<S>, dim3((unsigned)w), dim3(256), 0, s, ui + a * tsz,
                     (const int32_t *)nullptr, (const double *)c.d_t2 + a * T, c.d_t2 + a * T);
  if (mw) {
    unsigned long long *o = (unsigned long long *)c.d_s2 + a * T;
    hipLaunchKernelGGL(k_fill_sentinel, dim3(nblk(w * T)), dim3(256), 0, s, w * T, o);
    hipLaunchKernelGGL(k_bwd_sweep_mw<S>, dim3((unsigned)std::min<int64_t>(W, w)), dim3(SW), 0, s,
                       w, tb, sep, (const double *)ys, o, c.d_sweep_err);
    hipLaunchKernelGGL(k_collect, dim3(nblk(w * T)), dim3(256), 0, s, w * T,
                       (const unsigned long long *)o, ys);
  } else {
    hipLaunchKernelGGL(k_bwd_sweep<S>, dim3(1), dim3(SW), 0, s, w, tb, sep, ys);
  }
  const int64_t r0 = a - tb > 0 ? a - tb : 0, q0 = nb - tb > 0 ? nb - tb : 0;
  hipLaunchKernelGGL((k_rows_sub<S, false>), dim3((unsigned)(a - r0)), dim3(256), 0, s, tb, bt, r0,
                     a, rt, (const double *)c.d_t2, (int64_t)0, c.d_t2, (int64_t)0);
  hipLaunchKernelGGL((k_rows_sub<S, false>), dim3((unsigned)(nb - q0)), dim3(256), 0, s, tb, br, q0,
                     nb, nb + w, (const double *)c.d_t2, N, c.d_tw_dr, (int64_t)0);
  if (mw) {
    hipLaunchKernelGGL(k_fill_sentinel, dim3(nblk(a * T)), dim3(256), 0, s, a * T,
                       (unsigned long long *)c.d_s2);
    hipLaunchKernelGGL(k_fill_sentinel, dim3(nblk(nb * T)), dim3(256), 0, s, nb * T,
                       (unsigned long long *)c.d_tw_s2);
    hipLaunchKernelGGL((k_sweep2_mw<S, 1>), dim3(2 * W), dim3(SW), 0, s, tb, W,
                       SweepChain{a, bt, c.d_t2, c.d_s2}, SweepChain{nb, br, c.d_tw_dr, c.d_tw_s2},
                       c.d_sweep_err);
    hipLaunchKernelGGL(k_collect, dim3(nblk(a * T)), dim3(256), 0, s, a * T,
                       (const unsigned long long *)c.d_s2, d_x);
    hipLaunchKernelGGL(k_rev_copy<true>, dim3(nblk(nb * T)), dim3(256), 0, s, ns, nb * T, N,
                       (const double *)c.d_tw_s2, d_x);
  } else {
    hipLaunchKernelGGL((k_sweep2<S, 1>), dim3(2), dim3(SW), 0, s, tb,
                       SweepChain{a, bt, nullptr, c.d_t2}, SweepChain{nb, br, nullptr, c.d_tw_dr});
    MAFEM_HIP(hipMemcpyAsync(d_x, c.d_t2, sizeof(double) * a * T, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_rev_copy<false>, dim3(nblk(nb * T)), dim3(256), 0, s, ns, nb * T, N,
                       (const double *)c.d_tw_dr, d_x);
  }
  MAFEM_HIP(hipMemcpyAsync(d_x + a * T, c.d_t2 + a * T, sizeof(double) * w * T,
                           hipMemcpyDeviceToDevice, s));
  MAFEM_HIP(hipGetLastError());
}

template <class S>
static void solve_t(Ctx &c, const S *band, const S *li, const S *ui, const double *d_rhs,
                    double *d_x) {
  if (c.tw_a > 0) {
    solve_tw<S>(c, band, li, ui, d_rhs, d_x);
    return;
  }
  if (c.sweep_groups > 1) {
    // c = L_II^-1 P b -> t2; y (sentinel-filled s1) = forward sweep; d = U_II^-1 y -> t2;
    // x (sentinel-filled s2) = backward sweep, collected into d_x
    const int64_t nv = c.nt * T;
    const unsigned nb = (unsigned)((nv + 255) / 256);
    hipLaunchKernelGGL(k_diag_apply<S>, dim3((unsigned)c.nt), dim3(256), 0, c.stream, li, c.d_piv,
                       d_rhs, c.d_t2);
    hipLaunchKernelGGL(k_fill_sentinel, dim3(nb), dim3(256), 0, c.stream, nv,
                       (unsigned long long *)c.d_s1);
    hipLaunchKernelGGL(k_fwd_sweep_mw<S>, dim3(c.sweep_groups), dim3(SW), 0, c.stream, c.nt, c.tb,
                       band, c.d_t2, (unsigned long long *)c.d_s1, c.d_sweep_err);
    hipLaunchKernelGGL(k_diag_apply_atomic<S>, dim3((unsigned)c.nt), dim3(256), 0, c.stream, ui,
                       c.d_s1, c.d_t2);
    hipLaunchKernelGGL(k_fill_sentinel, dim3(nb), dim3(256), 0, c.stream, nv,
                       (unsigned long long *)c.d_s2);
    hipLaunchKernelGGL(k_bwd_sweep_mw<S>, dim3(c.sweep_groups), dim3(SW), 0, c.stream, c.nt, c.tb,
                       band, c.d_t2, (unsigned long long *)c.d_s2, c.d_sweep_err);
    hipLaunchKernelGGL(k_collect, dim3(nb), dim3(256), 0, c.stream, nv,
                       (const unsigned long long *)c.d_s2, d_x);
    MAFEM_HIP(hipGetLastError());
    return;
  }
  // y = L^-1 P b in place in d_x, then x = U^-1 y
  hipLaunchKernelGGL(k_diag_apply<S>, dim3((unsigned)c.nt), dim3(256), 0, c.stream, li, c.d_piv,
                     d_rhs, c.d_t2);
  hipLaunchKernelGGL(k_fwd_sweep<S>, dim3(1), dim3(SW), 0, c.stream, c.nt, c.tb, band, c.d_t2);
  hipLaunchKernelGGL(k_diag_apply<S>, dim3((unsigned)c.nt), dim3(256), 0, c.stream, ui,
                     (const int32_t *)nullptr, c.d_t2, d_x);
  hipLaunchKernelGGL(k_bwd_sweep<S>, dim3(1), dim3(SW), 0, c.stream, c.nt, c.tb, band, d_x);
  MAFEM_HIP(hipGetLastError());
}

// Allocates the band storage of the requested precision on first use.
static void ensure_band(Ctx &c, int prec) {
  const int64_t ntile = c.band_rows() * (2 * c.tb + 1) * (int64_t)T * T;
  const int64_t ninv = c.band_rows() * (int64_t)T * T;
  const size_t es = (prec == 32) ? sizeof(float) : sizeof(double);
  if (c.band_prec == prec && c.d_band) return;
  std::lock_guard<std::mutex> no_capture(capture_mutex());  // allocation: not during a capture
  c.batch_factor_prec = 0;
  c.have_factor = false;
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

void band_alloc(Ctx &c, int prec) { ensure_band(c, prec); }

// nested-dissection fronts (frontal.inc, below)
void nd_factor(Ctx &c, int prec, bool perturb);
void nd_solve(Ctx &c, const double *rhs, double *x);

void band_fill(Ctx &c, int prec) {
  if (c.ordering == 1) {  // the fronts gather the tangent level by level while factoring
    c.nd_fill_prec = prec;
    return;
  }
  ensure_band(c, prec);
  c.batch_factor_prec = 0;  // the band no longer holds a batch's factorization
  if (prec == 32)
    fill_t<float>(c, (float *)c.d_band);
  else
    fill_t<double>(c, (double *)c.d_band);
}

void band_factor(Ctx &c, bool perturb) {
  ++c.n_factor;
  if (c.ordering == 1) {
    nd_factor(c, c.nd_fill_prec, perturb);
    return;
  }
  if (c.band_prec == 32) {
    if (perturb) factor_run<float, true>(c, (float *)c.d_band, (float *)c.d_linv, (float *)c.d_uinv);
    else factor_run<float, false>(c, (float *)c.d_band, (float *)c.d_linv, (float *)c.d_uinv);
  } else {
    if (perturb)
      factor_run<double, true>(c, (double *)c.d_band, (double *)c.d_linv, (double *)c.d_uinv);
    else
      factor_run<double, false>(c, (double *)c.d_band, (double *)c.d_linv, (double *)c.d_uinv);
  }
}

void band_solve(Ctx &c, const double *d_rhs, double *d_x) {
  ++c.n_band_solve;
  if (c.ordering == 1) {
    nd_solve(c, d_rhs, d_x);
    return;
  }
  if (c.band_prec == 32)
    solve_t<float>(c, (const float *)c.d_band, (const float *)c.d_linv, (const float *)c.d_uinv,
                   d_rhs, d_x);
  else
    solve_t<double>(c, (const double *)c.d_band, (const double *)c.d_linv,
                    (const double *)c.d_uinv, d_rhs, d_x);
}

template <class S>
static void b_factor_t(BatchPlan &B, const std::vector<std::vector<StepMap>> &steps) {
  const BatchFlags &f = B.f;
  hipLaunchKernelGGL(kb_bsr_to_band<S>, dim3(B.bsr.nblk), dim3(256), 0, B.stream, B.bsr, B.d_pd,
                     f.on, f.fac);
  hipLaunchKernelGGL(kb_band_pad<S>, dim3(B.P), dim3(64), 0, B.stream, B.d_pd, f.on, f.fac);
  hipLaunchKernelGGL(kb_amax_zero<S>, dim3((unsigned)((B.P + 255) / 256)), dim3(256), 0,
                     B.stream, B.P, B.d_pd, f.on, f.fac, f.pert);
  hipLaunchKernelGGL(kb_amax<S>, dim3(B.bsr.nblk), dim3(256), 0, B.stream, B.bsr, B.d_pd, f.on,
                     f.fac, f.pert);
  MAFEM_HIP(hipGetLastError());
  for (int64_t K = 0; K < (int64_t)steps.size(); ++K)
    for (const StepMap &sm : steps[K]) {
      const int32_t nd = sm.blk0[sm.ncls];
      StepMap smp = sm, smu = sm;
      int32_t bp = 0, bu = 0;
      for (int c = 0; c < sm.ncls; ++c) {
        smp.blk0[c] = bp;
        smu.blk0[c] = bu;
        bp += sm.count[c] * 3 * sm.tb[c];
        bu += sm.count[c] * sm.tb[c] * sm.tb[c];
      }
      smp.blk0[sm.ncls] = bp;
      smu.blk0[sm.ncls] = bu;
      hipLaunchKernelGGL(kb_diag<S>, dim3(nd), dim3(256), 0, B.stream, sm, K, B.d_order, B.d_pd,
                         f.on, f.fac, f.pert);
      hipLaunchKernelGGL(kb_panel<S>, dim3(bp), dim3(64), 0, B.stream, smp, K, B.d_order, B.d_pd,
                         f.on, f.fac);
      if (bu > 0)
        hipLaunchKernelGGL(kb_update<S>, dim3(bu), dim3(256), 0, B.stream, smu, K, B.d_order,
                           B.d_pd, f.on, f.fac);
    }
  hipLaunchKernelGGL(kb_tile_inv<S>, dim3(B.tiles.nblk), dim3(64), 0, B.stream, B.tiles, B.d_pd,
                     f.on, f.fac);
  hipLaunchKernelGGL(kb_fold<S>, dim3(B.fold.nblk), dim3(256), 0, B.stream, B.fold, B.d_pd, f.on,
                     f.fac);
  MAFEM_HIP(hipGetLastError());
}

void b_factor(BatchPlan &B) {
  hipLaunchKernelGGL(kb_zero_band, dim3(B.zero.nblk), dim3(256), 0, B.stream, B.zero, B.d_pd,
                     B.f.on, B.f.fac);
  MAFEM_HIP(hipGetLastError());
  if (B.has32) b_factor_t<float>(B, B.steps32);
  if (B.has64) b_factor_t<double>(B, B.steps64);
}

template <class S>
static void b_band_solve_t(BatchPlan &B, const uint8_t *mask) {
  const bool mw = B.mw.nblk > 0;
  hipLaunchKernelGGL(kb_diag_apply<S>, dim3(B.tiles.nblk), dim3(256), 0, B.stream, B.tiles, B.d_pd,
                     mask, 0);
  hipLaunchKernelGGL(kb_fwd_sweep<S>, dim3(B.P), dim3(SW), 0, B.stream, B.d_pd, mask);
  if (mw) {
    hipLaunchKernelGGL(kb_fill_sentinel, dim3(B.vec.nblk), dim3(256), 0, B.stream, B.vec, B.d_pd,
                       mask, 0, prec_of<S>());
    hipLaunchKernelGGL(kb_sweep_mw<S>, dim3(B.mw.nblk), dim3(SW), 0, B.stream, B.mw, B.d_pd, mask, 0);
  }
  hipLaunchKernelGGL(kb_diag_apply<S>, dim3(B.tiles.nblk), dim3(256), 0, B.stream, B.tiles, B.d_pd,
                     mask, 1);
  hipLaunchKernelGGL(kb_bwd_sweep<S>, dim3(B.P), dim3(SW), 0, B.stream, B.d_pd, mask);
  if (mw) {
    hipLaunchKernelGGL(kb_fill_sentinel, dim3(B.vec.nblk), dim3(256), 0, B.stream, B.vec, B.d_pd,
                       mask, 1, prec_of<S>());
    hipLaunchKernelGGL(kb_sweep_mw<S>, dim3(B.mw.nblk), dim3(SW), 0, B.stream, B.mw, B.d_pd, mask, 1);
    hipLaunchKernelGGL(kb_collect, dim3(B.vec.nblk), dim3(256), 0, B.stream, B.vec, B.d_pd, mask,
                       prec_of<S>());
  }
  MAFEM_HIP(hipGetLastError());
}

void b_band_solve(BatchPlan &B, const uint8_t *mask) {
  if (B.has32) b_band_solve_t<float>(B, mask);
  if (B.has64) b_band_solve_t<double>(B, mask);
}

}  // namespace mafem

// The multifrontal factorization shares the tile bodies above.
#include "frontal.inc"
