// Dense complex LU with partial pivoting, triangular solves and
// mixed-precision GMRES iterative refinement on gfx950.
//
// Replaces the LAPACK calls of solve_multi_linsys (linalg.f90:159-180,
// ZGETRF + ZGETRS) behind solve_systems (solver.f90:325-357).
//
// Factorisation: blocked right-looking LU (outer block NB = 256) whose
// trailing update A22 -= L21 U12 runs on the matrix cores
// (v_mfma_f64_16x16x4f64 for complex128 with four real products per complex
// product; v_mfma_f32_32x32x2f32 for complex64 with three, Gauss's form),
// both fed by direct global->LDS transfers.
// The panel is factored in inner blocks of IB = 32 columns with one launch
// per column: every block reduces the previous launch's per-block pivot
// candidates (|re|+|im|, lowest index on ties, as LAPACK's IZAMAX), the
// block holding the pivot row swaps it with the diagonal row across the
// panel, every row below is scaled and updated in registers, and each block
// leaves its candidate for the next column.  The remaining panel columns
// then get a small triangular solve and an MFMA update.  Row interchanges
// reach the columns left and right of the panel through a laswp kernel, so
// the factors are in ZGETRF layout (unit L below the diagonal, U on and
// above, 1-based ipiv).
//
// Solves: the row permutation is applied as one gather; forward and back
// substitution work block by block with precomputed inverses of the NB x NB
// diagonal blocks (x_b = inv(T_bb) r_b, then a GEMV-style update of the
// rest), so every step is parallel.  The inverses are formed and stored in
// complex128 for both factor precisions: an inverse rounded to complex64
// carries the block's condition number times 6e-8 as error, which for the
// U blocks of large systems costs the FP32 preconditioner most of its
// accuracy.
//
// Mixed precision: the complex128 matrix Z stays resident; a copy scaled by
// power-of-two row and column factors (exact in binary) is rounded to
// complex64 and factored.  GMRES on the FP64 system, preconditioned by the
// FP32 factors, refines the solution until the complex128 backward error
// ||b - Zx||_inf / (||Z||_inf ||x||_inf + ||b||_inf) reaches the tolerance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "mom_internal.hpp"

namespace mom {

template <typename T> struct Cx;
// acc_row(q, v): row (0..15) of the 16x16 output held by accumulator v of a
// lane in row group q = lane/16 (column = lane%16).  The two instructions
// differ: f32 16x16x4 keeps 4 consecutive rows per lane, f64 16x16x4
// interleaves them (checked by mom_test_mfma_layout).
template <> struct Cx<float> {
  using type = float2;
  typedef float acc4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ int acc_row(int q, int v) { return 4 * q + v; }
};
template <> struct Cx<double> {
  using type = double2;
  typedef double acc4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ int acc_row(int q, int v) { return q + 4 * v; }
};

template <typename T> __device__ __forceinline__ typename Cx<T>::acc4 mfma4(T a, T b, typename Cx<T>::acc4 c);
template <> __device__ __forceinline__ Cx<float>::acc4 mfma4<float>(float a, float b, Cx<float>::acc4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <> __device__ __forceinline__ Cx<double>::acc4 mfma4<double>(double a, double b, Cx<double>::acc4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

constexpr int NB = 256;  // outer block (trailing update depth)
constexpr int IB = 32;   // inner panel block

// ---------------------------------------------------------------------------
// Trailing update C -= A B (all complex, column-major), on the matrix cores.
//  - workgroup of 4 waves in a 2 x 2 arrangement; wave tile (16 WM) x (16 WN)
//    complex, workgroup tile BM x BN = (32 WM) x (32 WN);
//  - K is walked in chunks of BK complex: the next chunk is fetched into
//    registers while the current one is multiplied out of LDS (two LDS
//    buffers, planar re/im), one barrier per chunk;
//  - LDS rows are padded by 16 elements so that the four k-rows a wave reads
//    per MFMA operand fall into disjoint bank groups;
//  - workgroups are dispatched to the 8 XCDs round-robin and each XCD has
//    its own L2, so the launch index is first remapped to make each XCD's
//    workgroups a contiguous range of tiles; within it tiles are visited in
//    groups of kGroupM row blocks so that concurrently running workgroups
//    of one XCD share A and B panels in their L2;
//  - M3 = false: four real products per complex product; M3 = true: three
//    (Gauss/3M: T1 = Ar Br, T2 = Ai Bi, T3 = (Ar+Ai)(Br+Bi)), the sums formed
//    in registers from the two staged planes.
constexpr int kGroupM = 8;
constexpr int kXcds = 8;

// Bijection launch index -> tile index under which XCD x (launch indices
// x, x+8, ...) receives one contiguous range of tile indices.
__device__ __forceinline__ int xcd_contiguous(int pid, int G) {
  const int per = (G + kXcds - 1) / kXcds;
  int tall = G % kXcds;
  if (tall == 0) tall = kXcds;
  const int x = pid % kXcds, l = pid / kXcds;
  return x < tall ? x * per + l : tall * per + (x - tall) * (per - 1) + l;
}

// C = beta C + sgn A B with sgn = +-1 and beta = 0 or 1 (beta0 != 0: C is
// only written); the factorisation's updates are sgn = -1, beta = 1.
template <typename T, int WM, int WN, int BK, bool M3>
__global__ __launch_bounds__(256) void gemm_sub_kernel(const typename Cx<T>::type *__restrict__ A, int64_t lda,
                                                       const typename Cx<T>::type *__restrict__ B, int64_t ldb,
                                                       typename Cx<T>::type *__restrict__ Cm, int64_t ldc, int M,
                                                       int N, int K, int gm, int gn, T sgn, int beta0) {
  using C2 = typename Cx<T>::type;
  using acc4 = typename Cx<T>::acc4;
  constexpr int BM = 32 * WM, BN = 32 * WN;
  constexpr int PA = BM + 16, PB = BN + 16;
  constexpr int NP = 2;  // planes: re, im (3M forms re+im in registers)
  constexpr int LA = BM * BK / 256, LB = BN * BK / 256;  // complex loads per thread
  static_assert(LA * 256 == BM * BK && LB * 256 == BN * BK, "tile");
  __shared__ T As[2][NP][BK][PA];
  __shared__ T Bs[2][NP][BK][PB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int wr = (wave & 1) * 16 * WM, wc = (wave >> 1) * 16 * WN;
  // grouped tile order
  const int pid = xcd_contiguous(blockIdx.x, gridDim.x);
  const int per_group = kGroupM * gn;
  const int first_m = (pid / per_group) * kGroupM;
  const int gsize = min(gm - first_m, kGroupM);
  const int pm = first_m + (pid % per_group) % gsize;
  const int pn = (pid % per_group) / gsize;
  const int64_t row0 = (int64_t)pm * BM, col0 = (int64_t)pn * BN;

  acc4 acc0[WM][WN], acc1[WM][WN], acc2[WM][WN];
#pragma unroll
  for (int i = 0; i < WM; i++)
#pragma unroll
    for (int j = 0; j < WN; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        acc0[i][j][v] = 0;
        acc1[i][j][v] = 0;
        if (M3) acc2[i][j][v] = 0;
      }
  C2 ra[LA], rb[LB];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < LA; u++) {
      const int idx = tid + 256 * u;
      const int r = idx % BM, k = idx / BM;
      const int64_t gr = row0 + r;
      const int gk = k0 + k;
      ra[u].x = 0;
      ra[u].y = 0;
      if (gr < M && gk < K) ra[u] = A[gr + (int64_t)gk * lda];
    }
#pragma unroll
    for (int u = 0; u < LB; u++) {
      const int idx = tid + 256 * u;
      const int k = idx % BK, c = idx / BK;
      const int64_t gc = col0 + c;
      const int gk = k0 + k;
      rb[u].x = 0;
      rb[u].y = 0;
      if (gc < N && gk < K) rb[u] = B[gk + gc * ldb];
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int u = 0; u < LA; u++) {
      const int idx = tid + 256 * u;
      const int r = idx % BM, k = idx / BM;
      As[buf][0][k][r] = ra[u].x;
      As[buf][1][k][r] = ra[u].y;
    }
#pragma unroll
    for (int u = 0; u < LB; u++) {
      const int idx = tid + 256 * u;
      const int k = idx % BK, c = idx / BK;
      Bs[buf][0][k][c] = rb[u].x;
      Bs[buf][1][k][c] = rb[u].y;
    }
  };
  fetch(0);
  stash(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += BK) {
    const bool more = k0 + BK < K;
    if (more) fetch(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T ar[WM], ai[WM], as_[WM], br[WN], bi[WN], bs[WN];
#pragma unroll
      for (int i = 0; i < WM; i++) {
        ar[i] = As[buf][0][kk + kq][wr + 16 * i + l16];
        ai[i] = As[buf][1][kk + kq][wr + 16 * i + l16];
        if (M3) as_[i] = ar[i] + ai[i];
      }
#pragma unroll
      for (int j = 0; j < WN; j++) {
        br[j] = Bs[buf][0][kk + kq][wc + 16 * j + l16];
        bi[j] = Bs[buf][1][kk + kq][wc + 16 * j + l16];
        if (M3) bs[j] = br[j] + bi[j];
      }
#pragma unroll
      for (int i = 0; i < WM; i++) {
        const T nai = -ai[i];
#pragma unroll
        for (int j = 0; j < WN; j++) {
          if (M3) {
            acc0[i][j] = mfma4<T>(ar[i], br[j], acc0[i][j]);
            acc1[i][j] = mfma4<T>(ai[i], bi[j], acc1[i][j]);
            acc2[i][j] = mfma4<T>(as_[i], bs[j], acc2[i][j]);
          } else {
            acc0[i][j] = mfma4<T>(ar[i], br[j], acc0[i][j]);
            acc0[i][j] = mfma4<T>(nai, bi[j], acc0[i][j]);
            acc1[i][j] = mfma4<T>(ar[i], bi[j], acc1[i][j]);
            acc1[i][j] = mfma4<T>(ai[i], br[j], acc1[i][j]);
          }
        }
      }
    }
    if (more) stash(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int i = 0; i < WM; i++)
#pragma unroll
    for (int j = 0; j < WN; j++) {
      const int64_t col = col0 + wc + 16 * j + l16;
      if (col >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int64_t row = row0 + wr + 16 * i + Cx<T>::acc_row(kq, v);
        if (row < M) {
          C2 *p = Cm + row + col * ldc;
          C2 c;
          if (beta0) {
            c.x = 0;
            c.y = 0;
          } else {
            c = *p;
          }
          if (M3) {
            c.x += sgn * (acc0[i][j][v] - acc1[i][j][v]);
            c.y += sgn * (acc2[i][j][v] - acc0[i][j][v] - acc1[i][j][v]);
          } else {
            c.x += sgn * acc0[i][j][v];
            c.y += sgn * acc1[i][j][v];
          }
          *p = c;
        }
      }
    }
}

// ---------------------------------------------------------------------------
// Trailing update C -= A B for complex64 with K = 32 or 256 (the inner-panel
// and outer-block depths), fed by direct global->LDS transfers:
//  - B (K x N, column-major) is first transposed into Bt[k][c] (row stride
//    ldbt, zero-padded to a multiple of the tile width), so that k-rows of
//    both operands are contiguous;
//  - each stage holds BK k-rows of the A tile (64 rows) and the B tile
//    (128 columns) as interleaved complex; every wave moves whole 1 KB
//    pieces with global_load_lds_dwordx4 (16 bytes per lane), three stages
//    in flight, one barrier per stage.  The K loop is fully unrolled, so
//    every stage buffer is a distinct LDS array and the only waits on the
//    transfers are the explicit vmcnt ones;
//  - 4 waves in 2 x 2, wave tile 32 x 64 complex = two 32x32 blocks of
//    v_mfma_f32_32x32x2f32 per product, three products per complex product
//    (3M), fragments of the next k-step read while the current MFMAs run;
//  - the update of C goes through LDS (idle stage buffers) so that lanes
//    read and write 16 contiguous bytes of a column instead of one complex
//    in 32 different columns.
// Rows of A beyond M are clamped to M - 2 (M even): they only feed output
// rows that are not stored.
#ifndef MOM_DMA_BK
#define MOM_DMA_BK 8
#endif
constexpr int kDmaBK = MOM_DMA_BK;  // K depth of one transfer stage of the 3M trailing update
#ifndef MOM_DMA_TM
#define MOM_DMA_TM 1
#endif
constexpr int kDmaTM = MOM_DMA_TM;  // 32-row MFMA tiles per wave along M (1: 32x64 wave tile, 2: 64x64)
constexpr int kDmaBM = 64 * kDmaTM, kDmaBN = 128;

template <typename C2>
__global__ __launch_bounds__(256) void transpose_panel_kernel(const C2 *__restrict__ B, int64_t ldb, int K, int N,
                                                              C2 *__restrict__ Bt, int64_t ldbt) {
  __shared__ C2 t[32][33];
  const int c0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const int c = c0 + r, k = k0 + tx;
    C2 z;
    z.x = 0;
    z.y = 0;
    t[r][tx] = (c < N && k < K) ? B[k + (int64_t)c * ldb] : z;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int k = k0 + r, c = c0 + tx;
    if (k < K && c < ldbt) Bt[(int64_t)k * ldbt + c] = t[tx][r];
  }
}

// s_waitcnt vmcnt(N) lgkmcnt(0) (expcnt unconstrained).  Before a stage
// barrier both are needed: vmcnt for this wave's transfers into the stage
// about to be read, lgkmcnt(0) so that this wave's reads of the stage about
// to be refilled have completed (the compiler may otherwise leave the last
// fragment reads in flight across the barrier, and another wave's transfer
// would overwrite them).
template <int N>
__device__ __forceinline__ void wait_stage() {
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (0x7 << 4));
}

#define MOM_LDS(p) ((__attribute__((address_space(3))) void *)(p))
#define MOM_COMPILER_FENCE() asm volatile("" ::: "memory")

// Role only names the launch site in profiles (0: updates, 1: the product
// that forms a U block from the inverse of the diagonal L block).
template <int NK, int Role = 0>
__global__ __launch_bounds__(256) void gemm3m_dma_kernel(const float2 *__restrict__ A, int64_t lda,
                                                         const float2 *__restrict__ Bt, int64_t ldbt,
                                                         float2 *__restrict__ Cm, int64_t ldc, int M, int N, int gm,
                                                         int gn) {
  constexpr int BK = kDmaBK, BM = kDmaBM, BN = kDmaBN, TM = kDmaTM, TN = 2;
  constexpr int NA = BK * BM / 128, NB = BK * BN / 128;  // 1 KB transfers per stage
  static_assert(NA % 4 == 0 && NB % 4 == 0, "transfers per stage must split over 4 waves");
  constexpr int PER_WAVE = NA / 4 + NB / 4;
  typedef float acc16 __attribute__((ext_vector_type(16)));
  __shared__ float2 As0[BK][BM], As1[BK][BM], As2[BK][BM];
  __shared__ float2 Bs0[BK][BN], Bs1[BK][BN], Bs2[BK][BN];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l32 = lane & 31, kh = lane >> 5;
  const int wr = (wave & 1) * 32 * TM, wc = (wave >> 1) * 32 * TN;
  const int pid = xcd_contiguous(blockIdx.x, gridDim.x);
  const int per_group = kGroupM * gn;
  const int first_m = (pid / per_group) * kGroupM;
  const int gsize = min(gm - first_m, kGroupM);
  const int pm = first_m + (pid % per_group) % gsize;
  const int pn = (pid % per_group) / gsize;
  const int64_t row0 = (int64_t)pm * BM, col0 = (int64_t)pn * BN;
  // transfer t of this wave covers stage element e = 128 (wave + 4 t) + 2 lane
  const float2 *pa[NA / 4];
  const float2 *pb[NB / 4];
#pragma unroll
  for (int t = 0; t < NA / 4; t++) {
    const int e = 128 * (wave + 4 * t) + 2 * lane, k = e / BM, r = e % BM;
    int64_t gr = row0 + r;
    if (gr > M - 2) gr = M - 2;
    pa[t] = A + gr + (int64_t)k * lda;
  }
#pragma unroll
  for (int t = 0; t < NB / 4; t++) {
    const int e = 128 * (wave + 4 * t) + 2 * lane, k = e / BN, c = e % BN;
    pb[t] = Bt + col0 + c + (int64_t)k * ldbt;
  }
  const int64_t astep = (int64_t)BK * lda, bstep = (int64_t)BK * ldbt;
  auto issue = [&](int c, float2(*as)[BM], float2(*bs)[BN]) {
#pragma unroll
    for (int t = 0; t < NA / 4; t++)
      __builtin_amdgcn_global_load_lds((const void *)(pa[t] + c * astep), MOM_LDS(&as[0][0] + 128 * (wave + 4 * t)),
                                       16, 0, 0);
#pragma unroll
    for (int t = 0; t < NB / 4; t++)
      __builtin_amdgcn_global_load_lds((const void *)(pb[t] + c * bstep), MOM_LDS(&bs[0][0] + 128 * (wave + 4 * t)),
                                       16, 0, 0);
  };
  auto stageA = [&](int st) -> float2(*)[BM] { return st == 0 ? As0 : (st == 1 ? As1 : As2); };
  auto stageB = [&](int st) -> float2(*)[BN] { return st == 0 ? Bs0 : (st == 1 ? Bs1 : Bs2); };
  acc16 acc0[TM][TN], acc1[TM][TN], acc2[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int v = 0; v < 16; v++) {
        acc0[i][j][v] = 0.f;
        acc1[i][j][v] = 0.f;
        acc2[i][j][v] = 0.f;
      }
  issue(0, As0, Bs0);
  if (NK > 1) issue(1, As1, Bs1);
#pragma unroll
  for (int it = 0; it < NK; it++) {
    // this wave's transfers of stage `it` have landed and its reads of the
    // previous stage are complete; after the barrier the same holds for
    // every wave, so stage `it` is readable and stage it + 2 (= it - 1)
    // can be refilled
    if (it + 1 < NK)
      wait_stage<PER_WAVE>();
    else
      wait_stage<0>();
    MOM_COMPILER_FENCE();
    __builtin_amdgcn_s_barrier();
    MOM_COMPILER_FENCE();
    // every fragment of this stage is read before the next transfers are
    // issued: the compiler makes an LDS read that follows a direct transfer
    // wait for that transfer
    float2(*as)[BM] = stageA(it % 3);
    float2(*bs)[BN] = stageB(it % 3);
    float2 fa[BK / 2][TM], fb[BK / 2][TN];
#pragma unroll
    for (int q = 0; q < BK / 2; q++) {
#pragma unroll
      for (int i = 0; i < TM; i++) fa[q][i] = as[2 * q + kh][wr + 32 * i + l32];
#pragma unroll
      for (int j = 0; j < TN; j++) fb[q][j] = bs[2 * q + kh][wc + 32 * j + l32];
    }
    if (it + 2 < NK) issue(it + 2, stageA((it + 2) % 3), stageB((it + 2) % 3));
#pragma unroll
    for (int kk = 0; kk < BK; kk += 2) {
      const int cur = kk / 2;
      float aS[TM], bS[TN];
#pragma unroll
      for (int i = 0; i < TM; i++) aS[i] = fa[cur][i].x + fa[cur][i].y;
#pragma unroll
      for (int j = 0; j < TN; j++) bS[j] = fb[cur][j].x + fb[cur][j].y;
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) {
          acc0[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[cur][i].x, fb[cur][j].x, acc0[i][j], 0, 0, 0);
          acc1[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[cur][i].y, fb[cur][j].y, acc1[i][j], 0, 0, 0);
          acc2[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(aS[i], bS[j], acc2[i][j], 0, 0, 0);
        }
    }
  }
  // 32x32x2 output: accumulator v of lane l is row 8 (v / 4) + 4 (l / 32) +
  // v % 4, column l % 32.  Each wave stages CW columns of a 32x32 block at a
  // time in its quarter of As0 ([col][row]), then lanes update pairs of rows.
  constexpr int PERW = BK * BM / 4;  // complex per wave
  constexpr int CW = PERW / 32;      // columns per chunk
  static_assert(CW >= 2 && 32 % CW == 0, "epilogue chunk");
  __syncthreads();  // every wave is done with the stage buffers
  float2 *stg = &As0[0][0] + wave * PERW;
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int cb = 0; cb < 32; cb += CW) {
        if (l32 >= cb && l32 < cb + CW) {
#pragma unroll
          for (int v = 0; v < 16; v++) {
            const int rr = 8 * (v / 4) + 4 * kh + (v % 4);
            stg[(l32 - cb) * 32 + rr] =
                make_float2(acc0[i][j][v] - acc1[i][j][v], acc2[i][j][v] - acc0[i][j][v] - acc1[i][j][v]);
          }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int e = lane; e < CW * 16; e += 64) {  // CW columns x 16 row pairs
          const int cc = e / 16, rr = 2 * (e % 16);
          const int64_t col = col0 + wc + 32 * j + cb + cc, row = row0 + wr + 32 * i + rr;
          const float4 d = *reinterpret_cast<const float4 *>(&stg[cc * 32 + rr]);
          if (col < N && row + 1 < M) {
            float4 *p = reinterpret_cast<float4 *>(Cm + row + col * ldc);
            float4 c = *p;
            c.x -= d.x;
            c.y -= d.y;
            c.z -= d.z;
            c.w -= d.w;
            *p = c;
          } else if (col < N && row < M) {
            float2 *p = Cm + row + col * ldc;
            float2 c = *p;
            c.x -= d.x;
            c.y -= d.y;
            *p = c;
          }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
}

// ---------------------------------------------------------------------------
// Trailing update C -= A B for complex128 with K = 32 or 256 on the FP64
// matrix cores, fed by direct global->LDS transfers like the complex64
// kernel above:
//  - B (K x N, column-major) is first transposed into Bt[k][c]
//    (transpose_panel_kernel), so that k-rows of both operands are
//    contiguous;
//  - each stage holds BK = 8 k-rows of the A tile (128 rows) and of the B
//    tile (64 columns) as interleaved complex; a transfer moves 1 KB with one
//    complex (16 bytes) per lane; three stages in flight, one barrier per
//    stage; the K loop is fully unrolled (explicit vmcnt waits only);
//  - 4 waves in 2 x 2, wave tile 64 x 32 complex = 4 x 2 blocks of
//    v_mfma_f64_16x16x4f64; fragments are whole complex (one 16-byte LDS
//    read gives the real and imaginary operand of a lane), read for the
//    next k-step while the current MFMAs run;
//  - M3 = false: four real products per complex product (the operation
//    count of ZGEMM); M3 = true: three (Gauss), with the sums formed in
//    registers;
//  - the MFMAs of one k-step are issued product by product over all eight
//    blocks, so that consecutive instructions never feed the same
//    accumulator;
//  - C is updated in place from the accumulators: lane l holds rows
//    l/16 + 4v (v = 0..3) of column l%16 of each 16x16 block.
// Rows of A beyond M are clamped to M - 1: they only feed output rows that
// are not stored.
#ifndef MOM_Z_TM
#define MOM_Z_TM 4
#define MOM_Z_TN 2
#endif
constexpr int kZTM = MOM_Z_TM, kZTN = MOM_Z_TN;  // 16x16 blocks per wave along M and N
constexpr int kZBK = 8, kZBM = 32 * kZTM, kZBN = 32 * kZTN;

template <int NK, bool M3>
__global__ __launch_bounds__(256) void zgemm_dma_kernel(const double2 *__restrict__ A, int64_t lda,
                                                        const double2 *__restrict__ Bt, int64_t ldbt,
                                                        double2 *__restrict__ Cm, int64_t ldc, int M, int N, int gm,
                                                        int gn) {
  constexpr int BK = kZBK, BM = kZBM, BN = kZBN, TM = kZTM, TN = kZTN;
  constexpr int NA = BK * BM / 64, NBT = BK * BN / 64;  // 1 KB transfers per stage
  static_assert(NA % 4 == 0 && NBT % 4 == 0, "transfers per stage must split over 4 waves");
  constexpr int PER_WAVE = NA / 4 + NBT / 4;
  typedef double acc4 __attribute__((ext_vector_type(4)));
  __shared__ double2 As0[BK][BM], As1[BK][BM], As2[BK][BM];
  __shared__ double2 Bs0[BK][BN], Bs1[BK][BN], Bs2[BK][BN];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int wr = (wave & 1) * 16 * TM, wc = (wave >> 1) * 16 * TN;
  const int pid = xcd_contiguous(blockIdx.x, gridDim.x);
  const int per_group = kGroupM * gn;
  const int first_m = (pid / per_group) * kGroupM;
  const int gsize = min(gm - first_m, kGroupM);
  const int pm = first_m + (pid % per_group) % gsize;
  const int pn = (pid % per_group) / gsize;
  const int64_t row0 = (int64_t)pm * BM, col0 = (int64_t)pn * BN;
  // transfer t of this wave covers stage element e = 64 (wave + 4 t) + lane
  const double2 *pa[NA / 4];
  const double2 *pb[NBT / 4];
#pragma unroll
  for (int t = 0; t < NA / 4; t++) {
    const int e = 64 * (wave + 4 * t) + lane, k = e / BM, r = e % BM;
    int64_t gr = row0 + r;
    if (gr > M - 1) gr = M - 1;
    pa[t] = A + gr + (int64_t)k * lda;
  }
#pragma unroll
  for (int t = 0; t < NBT / 4; t++) {
    const int e = 64 * (wave + 4 * t) + lane, k = e / BN, c = e % BN;
    pb[t] = Bt + col0 + c + (int64_t)k * ldbt;
  }
  const int64_t astep = (int64_t)BK * lda, bstep = (int64_t)BK * ldbt;
  auto issue = [&](int c, double2(*as)[BM], double2(*bs)[BN]) {
#pragma unroll
    for (int t = 0; t < NA / 4; t++)
      __builtin_amdgcn_global_load_lds((const void *)(pa[t] + c * astep), MOM_LDS(&as[0][0] + 64 * (wave + 4 * t)),
                                       16, 0, 0);
#pragma unroll
    for (int t = 0; t < NBT / 4; t++)
      __builtin_amdgcn_global_load_lds((const void *)(pb[t] + c * bstep), MOM_LDS(&bs[0][0] + 64 * (wave + 4 * t)),
                                       16, 0, 0);
  };
  auto stageA = [&](int st) -> double2(*)[BM] { return st == 0 ? As0 : (st == 1 ? As1 : As2); };
  auto stageB = [&](int st) -> double2(*)[BN] { return st == 0 ? Bs0 : (st == 1 ? Bs1 : Bs2); };
  constexpr int NACC = M3 ? 3 : 2;
  acc4 acc[NACC][TM][TN];
#pragma unroll
  for (int a = 0; a < NACC; a++)
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++)
#pragma unroll
        for (int v = 0; v < 4; v++) acc[a][i][j][v] = 0.0;
  issue(0, As0, Bs0);
  if (NK > 1) issue(1, As1, Bs1);
#pragma unroll
  for (int it = 0; it < NK; it++) {
    // this wave's transfers of stage `it` have landed and its reads of the
    // previous stage are complete; after the barrier the same holds for
    // every wave, so stage `it` is readable and stage it + 2 (= it - 1)
    // can be refilled
    if (it + 1 < NK)
      wait_stage<PER_WAVE>();
    else
      wait_stage<0>();
    MOM_COMPILER_FENCE();
    __builtin_amdgcn_s_barrier();
    MOM_COMPILER_FENCE();
    // every fragment of this stage is read before the next transfers are
    // issued: the compiler makes an LDS read that follows a direct transfer
    // wait for that transfer
    double2(*as)[BM] = stageA(it % 3);
    double2(*bs)[BN] = stageB(it % 3);
    double2 fa[BK / 4][TM], fb[BK / 4][TN];
#pragma unroll
    for (int q = 0; q < BK / 4; q++) {
#pragma unroll
      for (int i = 0; i < TM; i++) fa[q][i] = as[4 * q + kq][wr + 16 * i + l16];
#pragma unroll
      for (int j = 0; j < TN; j++) fb[q][j] = bs[4 * q + kq][wc + 16 * j + l16];
    }
    if (it + 2 < NK) issue(it + 2, stageA((it + 2) % 3), stageB((it + 2) % 3));
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int cur = kk / 4;
      if (M3) {
        double aS[TM], bS[TN];
#pragma unroll
        for (int i = 0; i < TM; i++) aS[i] = fa[cur][i].x + fa[cur][i].y;
#pragma unroll
        for (int j = 0; j < TN; j++) bS[j] = fb[cur][j].x + fb[cur][j].y;
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++)
            acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[cur][i].x, fb[cur][j].x, acc[0][i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++)
            acc[1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[cur][i].y, fb[cur][j].y, acc[1][i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++)
            acc[NACC - 1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(aS[i], bS[j], acc[NACC - 1][i][j], 0, 0, 0);
      } else {
        double nai[TM];
#pragma unroll
        for (int i = 0; i < TM; i++) nai[i] = -fa[cur][i].y;
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++)
            acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[cur][i].x, fb[cur][j].x, acc[0][i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++)
            acc[1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[cur][i].x, fb[cur][j].y, acc[1][i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++)
            acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[i], fb[cur][j].y, acc[0][i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++)
            acc[1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[cur][i].y, fb[cur][j].x, acc[1][i][j], 0, 0, 0);
      }
    }
  }
  auto result = [&](int i, int j, int v) -> double2 {
    if (M3)
      return make_double2(acc[0][i][j][v] - acc[1][i][j][v],
                          acc[NACC - 1][i][j][v] - acc[0][i][j][v] - acc[1][i][j][v]);
    return make_double2(acc[0][i][j][v], acc[1][i][j][v]);
  };
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) {
      const int64_t col = col0 + wc + 16 * j + l16;
      if (col >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int64_t row = row0 + wr + 16 * i + kq + 4 * v;
        if (row < M) {
          double2 *p = Cm + row + col * ldc;
          double2 c = *p;
          const double2 r = result(i, j, v);
          c.x -= r.x;
          c.y -= r.y;
          *p = c;
        }
      }
    }
}

#ifndef MOM_Z_M3
#define MOM_Z_M3 false
#endif
constexpr bool kZM3 = MOM_Z_M3;  // complex128 trailing update by three real products instead of four

// Tile configuration (wave tile 16 WM x 16 WN complex, K chunk BK); the
// MOM_GEMM_* macros let a build choose another configuration.
#ifndef MOM_GEMM_F32_WM
#define MOM_GEMM_F32_WM 2
#define MOM_GEMM_F32_WN 2
#define MOM_GEMM_F32_BK 16
#define MOM_GEMM_F32_M3 true
#endif
template <typename T> struct GemmCfg;
template <> struct GemmCfg<float> {
  static constexpr int WM = MOM_GEMM_F32_WM, WN = MOM_GEMM_F32_WN, BK = MOM_GEMM_F32_BK;
  static constexpr bool M3 = MOM_GEMM_F32_M3;
};
template <> struct GemmCfg<double> {
  static constexpr int WM = 2, WN = 2, BK = 8;
  static constexpr bool M3 = false;
};

// Event-pair timing of the factorisation's kernel groups (enabled per
// system): trailing-update GEMM launches (with their algorithmic flops),
// the panel column kernels of each inner block, triangular solves and row
// interchanges.  Level 1 times the depth-NB trailing updates only (one pair
// per block column, no host wait inside a factorisation of up to kRing
// block columns); level 2 times every group.  Pairs live in a fixed ring;
// when it is full the host waits for the oldest pairs and folds them into
// the totals, so the number of live events stays bounded.
enum Phase { kPhGemm = 0, kPhPanel, kPhTrsm, kPhLaswp, kPhGemmNarrow, kPhases };

struct PhaseClock {
  static constexpr int kRing = 1024;  // event pairs
  bool enabled = false;
  bool all_phases = false;
  std::vector<hipEvent_t> pool;
  int used = 0;
  int phase_of[kRing];
  double flops_of[kRing];
  double seconds[kPhases] = {}, launches[kPhases] = {}, flops[kPhases] = {};
  void reset() {
    used = 0;
    for (int p = 0; p < kPhases; p++) seconds[p] = launches[p] = flops[p] = 0;
  }
  int begin(hipStream_t s) {
    if (pool.empty()) {
      pool.resize(2 * kRing);
      for (auto &e : pool) MOM_HIP(hipEventCreate(&e));
    }
    if (used == kRing) {
      int rc = collect();
      if (rc) return rc;
    }
    MOM_HIP(hipEventRecord(pool[2 * used], s));
    return 0;
  }
  int end(hipStream_t s, Phase ph, double fl = 0) {
    MOM_HIP(hipEventRecord(pool[2 * used + 1], s));
    phase_of[used] = ph;
    flops_of[used] = fl;
    used++;
    return 0;
  }
  // waits for the recorded pairs and adds them to the totals
  int collect() {
    if (used == 0) return 0;
    for (int i = 0; i < used; i++) {
      float ms = 0;
      MOM_HIP(hipEventSynchronize(pool[2 * i + 1]));  // pairs may sit on either stream
      MOM_HIP(hipEventElapsedTime(&ms, pool[2 * i], pool[2 * i + 1]));
      seconds[phase_of[i]] += ms * 1e-3;
      launches[phase_of[i]] += 1;
      flops[phase_of[i]] += flops_of[i];
    }
    used = 0;
    return 0;
  }
  ~PhaseClock() {
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

// Brackets a group of launches on stream s with an event pair when the
// clock is enabled.
struct Timed {
  PhaseClock *c;
  hipStream_t s;
  Phase ph;
  double fl;
  int rc = 0;
  Timed(PhaseClock *c_, hipStream_t s_, Phase ph_, double fl_ = 0)
      : c(c_ && c_->enabled && (c_->all_phases || ph_ == kPhGemm) ? c_ : nullptr), s(s_), ph(ph_), fl(fl_) {
    if (c) rc = c->begin(s);
  }
  int done() { return (c && !rc) ? c->end(s, ph, fl) : rc; }
};

// A stream of the factorisation with the scratch its GEMMs use (the
// transposed B panel of the direct-transfer kernels).  Two lanes never
// share scratch, so their launches may overlap.
struct Lane {
  hipStream_t s;
  DevBuf *bt;
};

static Lane main_lane(mom_ctx *ctx) { return Lane{ctx->stream, &ctx->panel_t}; }

// The look-ahead lane: second stream (default priority) and its events.
static int aux_lane(mom_ctx *ctx, Lane *l) {
  if (!ctx->aux) {
    MOM_HIP(hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
    MOM_HIP(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
    MOM_HIP(hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming));
  }
  *l = Lane{ctx->aux, &ctx->panel_t2};
  return 0;
}

// Whether the direct-transfer kernels take C -= A B of these shapes (they
// read B through a transposed copy, so B may alias C).  complex64 moves rows
// in pairs (8-byte elements, 16-byte transfers).
static bool dma_gemm_ok(int64_t M, int64_t K, int64_t lda, int64_t ldc, const void *A, const void *C) {
  return (K == 256 || K == 32) && M >= 2 && M % 2 == 0 && lda % 2 == 0 && ldc % 2 == 0 &&
         (reinterpret_cast<uintptr_t>(A) & 15) == 0 && (reinterpret_cast<uintptr_t>(C) & 15) == 0;
}
static bool zdma_gemm_ok(int64_t M, int64_t K, const void *A, const void *B, const void *C) {
  return (K == 256 || K == 32) && M >= 1 && (reinterpret_cast<uintptr_t>(A) & 15) == 0 &&
         (reinterpret_cast<uintptr_t>(B) & 15) == 0 && (reinterpret_cast<uintptr_t>(C) & 15) == 0;
}

// B (K x N, ldb) transposed into the lane's buffer Bt[k][c], padded with
// zero columns to a multiple of bn; returns Bt and its row stride.
template <typename C2>
static int transpose_b(const Lane &L, const C2 *B, int64_t ldb, int64_t K, int64_t N, int64_t bn, C2 **Bt,
                       int64_t *ldbt, PhaseClock *st) {
  *ldbt = (N + bn - 1) / bn * bn;
  const size_t bytes = sizeof(C2) * (size_t)K * *ldbt;
  if (L.bt->bytes < bytes) MOM_TRY(L.bt->alloc(bytes));
  *Bt = L.bt->as<C2>();
  Timed tt(st, L.s, kPhGemmNarrow, 0.0);
  hipLaunchKernelGGL(transpose_panel_kernel<C2>, dim3((unsigned)(*ldbt / 32), (unsigned)((K + 31) / 32)), dim3(256), 0,
                     L.s, B, ldb, (int)K, (int)N, *Bt, *ldbt);
  MOM_HIP(hipGetLastError());
  return tt.done();
}

// C = beta C + sgn A B (sgn = +-1, beta = 0 or 1) by the register-staged
// kernel, any shape; grids of fewer than kSmallGrid tiles take 32 x 32
// tiles instead.
constexpr int64_t kSmallGrid = 128;
template <typename T>
static int gemm_generic(hipStream_t s, const typename Cx<T>::type *A, int64_t lda, const typename Cx<T>::type *B,
                        int64_t ldb, typename Cx<T>::type *C, int64_t ldc, int64_t M, int64_t N, int64_t K, T sgn,
                        bool beta0) {
  constexpr int WM = GemmCfg<T>::WM, WN = GemmCfg<T>::WN;
  const int gm = (int)((M + 32 * WM - 1) / (32 * WM)), gn = (int)((N + 32 * WN - 1) / (32 * WN));
  if ((int64_t)gm * gn < kSmallGrid) {
    // few tiles (the 256 x 256 diagonal-block products of the block
    // solves): 32 x 32 tiles, four times the workgroups
    const int gm1 = (int)((M + 31) / 32), gn1 = (int)((N + 31) / 32);
    hipLaunchKernelGGL((gemm_sub_kernel<T, 1, 1, GemmCfg<T>::BK, GemmCfg<T>::M3>), dim3((unsigned)gm1 * gn1),
                       dim3(256), 0, s, A, lda, B, ldb, C, ldc, (int)M, (int)N, (int)K, gm1, gn1, sgn, beta0 ? 1 : 0);
  } else {
    hipLaunchKernelGGL((gemm_sub_kernel<T, WM, WN, GemmCfg<T>::BK, GemmCfg<T>::M3>), dim3((unsigned)gm * gn),
                       dim3(256), 0, s, A, lda, B, ldb, C, ldc, (int)M, (int)N, (int)K, gm, gn, sgn, beta0 ? 1 : 0);
  }
  MOM_HIP(hipGetLastError());
  return 0;
}

// C -= A B on lane L, all by the hand-written matrix-core kernels:
// complex64 by the 3M direct-transfer kernel for K = 32 and 256, complex128
// by the FP64 direct-transfer kernel for K = 32 and 256, other depths by
// the register-staged kernel.
// u_solve: the product forms a U block (U12 = A12 - E A12, E from
// unit_lower_inverse_kernel); it is timed with the triangular solves and
// launched as its own kernel instance so profiles keep it apart from the
// updates.
template <typename T>
static int gemm_sub(const Lane &L, const typename Cx<T>::type *A, int64_t lda, const typename Cx<T>::type *B,
                    int64_t ldb, typename Cx<T>::type *C, int64_t ldc, int64_t M, int64_t N, int64_t K,
                    PhaseClock *st = nullptr, bool u_solve = false) {
  if (M <= 0 || N <= 0 || K <= 0) return 0;
  hipStream_t s = L.s;
  // updates of depth NB are the trailing updates; shallower ones (inner
  // panel, triangular-solve blocks) are timed separately.  The bracket holds
  // the GEMM kernel alone (the panel transpose is timed with the narrow
  // group of its own).
  const Phase ph = u_solve ? kPhTrsm : K >= NB ? kPhGemm : kPhGemmNarrow;
  const double fl = u_solve ? 0.0 : 8.0 * (double)M * (double)N * (double)K;
  if constexpr (std::is_same<T, double>::value) {
    if (u_solve) return fail(MOM_ERR_ARG, "U-block product is a complex64 path");
    if (zdma_gemm_ok(M, K, A, B, C)) {
      const int gm = (int)((M + kZBM - 1) / kZBM), gn = (int)((N + kZBN - 1) / kZBN);
      double2 *Bt;
      int64_t ldbt;
      MOM_TRY(transpose_b<double2>(L, B, ldb, K, N, kZBN, &Bt, &ldbt, st));
      Timed t(st, s, ph, fl);
      if (K == 256)
        hipLaunchKernelGGL((zgemm_dma_kernel<256 / kZBK, kZM3>), dim3((unsigned)gm * gn), dim3(256), 0, s, A, lda, Bt,
                           ldbt, C, ldc, (int)M, (int)N, gm, gn);
      else
        hipLaunchKernelGGL((zgemm_dma_kernel<32 / kZBK, kZM3>), dim3((unsigned)gm * gn), dim3(256), 0, s, A, lda, Bt,
                           ldbt, C, ldc, (int)M, (int)N, gm, gn);
      MOM_HIP(hipGetLastError());
      return t.done();
    }
    Timed t(st, s, ph, fl);
    MOM_TRY(gemm_generic<double>(s, A, lda, B, ldb, C, ldc, M, N, K, -1.0, false));
    return t.done();
  } else {
    const bool dma = dma_gemm_ok(M, K, lda, ldc, A, C);
    if (u_solve && (!dma || K != NB)) return fail(MOM_ERR_ARG, "U-block product needs the depth-NB direct-transfer kernel");
    if (dma) {
      const int gm = (int)((M + kDmaBM - 1) / kDmaBM), gn = (int)((N + kDmaBN - 1) / kDmaBN);
      float2 *Bt;
      int64_t ldbt;
      MOM_TRY(transpose_b<float2>(L, B, ldb, K, N, kDmaBN, &Bt, &ldbt, st));
      Timed t(st, s, ph, fl);
      if (K == 256 && u_solve)
        hipLaunchKernelGGL((gemm3m_dma_kernel<256 / kDmaBK, 1>), dim3((unsigned)gm * gn), dim3(256), 0, s, A, lda, Bt,
                           ldbt, C, ldc, (int)M, (int)N, gm, gn);
      else if (K == 256)
        hipLaunchKernelGGL(gemm3m_dma_kernel<256 / kDmaBK>, dim3((unsigned)gm * gn), dim3(256), 0, s, A, lda, Bt, ldbt,
                           C, ldc, (int)M, (int)N, gm, gn);
      else
        hipLaunchKernelGGL(gemm3m_dma_kernel<32 / kDmaBK>, dim3((unsigned)gm * gn), dim3(256), 0, s, A, lda, Bt, ldbt,
                           C, ldc, (int)M, (int)N, gm, gn);
      MOM_HIP(hipGetLastError());
      return t.done();
    }
    Timed t(st, s, ph, fl);
    MOM_TRY(gemm_generic<float>(s, A, lda, B, ldb, C, ldc, M, N, K, -1.0f, false));
    return t.done();
  }
}

// C = beta C + sgn A B in complex128 for the block solves and residuals
// (sgn = +-1, beta = 0 or 1).  C -= A B with K = NB goes to the
// direct-transfer kernel; a deep product C -= A B (K > NB, the residual
// B - Z X) runs as depth-NB slices through it, and the rest through the
// register-staged kernel.
static int zgemm(const Lane &L, int64_t M, int64_t N, int64_t K, double sgn, bool beta0, const double2 *A,
                 int64_t lda, const double2 *B, int64_t ldb, double2 *C, int64_t ldc) {
  if (M <= 0 || N <= 0) return 0;
  if (K <= 0) {
    if (!beta0) return 0;
    return gemm_generic<double>(L.s, A, lda, B, ldb, C, ldc, M, N, 0, sgn, true);
  }
  if (sgn < 0 && !beta0 && K >= NB && zdma_gemm_ok(M, NB, A, B, C)) {
    int64_t k0 = 0;
    for (; k0 + NB <= K; k0 += NB) MOM_TRY(gemm_sub<double>(L, A + k0 * lda, lda, B + k0, ldb, C, ldc, M, N, NB));
    if (k0 < K) MOM_TRY(gemm_generic<double>(L.s, A + k0 * lda, lda, B + k0, ldb, C, ldc, M, N, K - k0, -1.0, false));
    return 0;
  }
  return gemm_generic<double>(L.s, A, lda, B, ldb, C, ldc, M, N, K, sgn, beta0);
}

// ---------------------------------------------------------------------------
// Panel kernels.
// Pivot candidates of one panel column, one per block of 256 rows: the
// block's largest |re|+|im| (lowest row on ties) and that row's values in
// the inner block's columns.  Two parities: the kernel for column j reads
// parity j&1 and writes the candidates of column j+1 into the other one, so
// blocks exchange pivot data only across kernel boundaries (no device-scope
// fences inside a kernel).
struct PivotWork {
  double *val;   // [2][mb]
  int *idx;      // [2][mb]
  void *row;     // [2][mb][32] complex of the factor's type
  int mb;        // slots per parity (>= blocks of the widest launch)
  int *info;     // first zero pivot (1-based), 0 if none
  int *plan;     // laswp plan of one panel: 1 + 4 NB ints
  int *tsel;     // tournament candidate sets: [2][mb][IB] row indices
  float2 *lu_top;  // tournament: L11 U11 of the inner block by rows, IB x IB
  int tournament;  // complex64 panels by tournament pivoting (the system's choice or the test hook)
};


template <typename T>
__device__ __forceinline__ T cabs1(typename Cx<T>::type v) {
  return fabs(v.x) + fabs(v.y);
}

__device__ __forceinline__ void argmax_merge(double &v, int &i, int &w, double v2, int i2, int w2) {
  if (v2 > v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
    w = w2;
  }
}

// Block-wide argmax of (value, row, tag) over 256 threads; every thread
// receives the result.  The order (value descending, row ascending) is
// total, so the result does not depend on the reduction tree.
__device__ __forceinline__ void block_argmax(double &v, int &i, int &w, double *sv, int *si, int *sw) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double v2 = __shfl_xor(v, o);
    const int i2 = __shfl_xor(i, o);
    const int w2 = __shfl_xor(w, o);
    argmax_merge(v, i, w, v2, i2, w2);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    sv[wv] = v;
    si[wv] = i;
    sw[wv] = w;
  }
  __syncthreads();
  v = sv[0];
  i = si[0];
  w = sw[0];
#pragma unroll
  for (int k = 1; k < 4; k++) argmax_merge(v, i, w, sv[k], si[k], sw[k]);
  __syncthreads();
}

// Candidates of column `col` over rows col..n-1 (block b: rows col+256b ..),
// with the candidate row's values in columns [col, cend) stored at the
// inner-block offsets c - i0.
template <typename T>
__global__ __launch_bounds__(256) void pivot_candidates_kernel(const typename Cx<T>::type *A, int64_t ld, int col,
                                                                int n, int i0, int cend, PivotWork w, int par) {
  using C2 = typename Cx<T>::type;
  __shared__ double sv[4];
  __shared__ int si[4], sw[4];
  const int tid = threadIdx.x;
  const int64_t r = (int64_t)col + (int64_t)blockIdx.x * 256 + tid;
  double v = -1.0;
  int idx = n, win = tid;
  if (r < n) {
    v = (double)cabs1<T>(A[(int64_t)col * ld + r]);
    idx = (int)r;
  }
  block_argmax(v, idx, win, sv, si, sw);
  // value and row index always (n: no candidate, e.g. all NaN), the row's
  // values only for a candidate
  if (tid == win) {
    const size_t slot = (size_t)par * w.mb + blockIdx.x;
    w.val[slot] = v;
    w.idx[slot] = idx;
  }
  if (tid == win && idx < n) {
    const size_t slot = (size_t)par * w.mb + blockIdx.x;
    C2 *row = (C2 *)w.row + slot * IB;
    C2 x[IB];
#pragma unroll
    for (int q = 0; q < IB; q++)
      if (i0 + q >= col && i0 + q < cend) x[q] = A[(int64_t)(i0 + q) * ld + r];
#pragma unroll
    for (int q = 0; q < IB; q++)
      if (i0 + q >= col && i0 + q < cend) row[q] = x[q];
  }
}

// One column step j of the inner block [i0, jend) of the panel [pc0, pc1):
//  - every block reduces the candidates of column j (parity j&1) to the
//    pivot row p and takes the pivot row's values from the candidate slot;
//  - the block holding row p moves row j's old values into row p and swaps
//    rows j and p in the panel columns outside the inner block, and writes
//    the pivot row into row j;
//  - every row i > j is scaled (l = a_ij / pivot) and gets the rank-1
//    update of columns j+1..jend-1 in registers;
//  - with fuse_next, each block stores its candidate for column j+1.
template <typename T>
__global__ __launch_bounds__(256) void panel_column_kernel(typename Cx<T>::type *A, int64_t ld, int j, int n, int i0,
                                                            int jend, int pc0, int pc1, int *ipiv, PivotWork w,
                                                            int nb_in, int fuse_next) {
  using C2 = typename Cx<T>::type;
  __shared__ C2 u[IB];
  __shared__ double sv[4];
  __shared__ int si[4], sw[4];
  const int tid = threadIdx.x, b = blockIdx.x;
  const int64_t i = (int64_t)j + 1 + (int64_t)b * 256 + tid;
  const bool own = i < n;
  const int k = j - i0, kend = jend - i0;
  C2 v[IB];
#pragma unroll
  for (int s = 0; s < IB; s++) {
    v[s].x = 0;
    v[s].y = 0;
    if (own && s >= k && s < kend) v[s] = A[(int64_t)(i0 + s) * ld + i];
  }
  // pivot of column j from the block candidates
  const int par = j & 1;
  double bv = -1.0;
  int bi = n, bw = 0;
  for (int t = tid; t < nb_in; t += 256)
    argmax_merge(bv, bi, bw, w.val[(size_t)par * w.mb + t], w.idx[(size_t)par * w.mb + t], t);
  block_argmax(bv, bi, bw, sv, si, sw);
  const bool none = !(bi >= j && bi < n);  // no candidate compares (NaN): keep row j
  const int p = none ? j : bi;
  const bool zero = !(bv > 0.0);
  const C2 *cand = (const C2 *)w.row + ((size_t)par * w.mb + bw) * IB;
  if (tid < kend - k) u[k + tid] = none ? A[(int64_t)(j + tid) * ld + j] : cand[k + tid];
  __syncthreads();
  if (b == 0 && tid == 0) {
    ipiv[j] = p + 1;
    if (zero) atomicCAS(w.info, 0, j + 1);
  }
  if (p > j && b == (int)((p - j - 1) / 256)) {
    if (i == p) {
#pragma unroll
      for (int s = 0; s < IB; s++)
        if (s >= k && s < kend) v[s] = A[(int64_t)(i0 + s) * ld + j];
    }
    for (int c = pc0 + tid; c < pc1; c += 256) {
      if (c >= j && c < jend) continue;
      C2 *cc = A + (int64_t)c * ld;
      const C2 x = cc[j], y = cc[p];
      cc[j] = y;
      cc[p] = x;
    }
    __syncthreads();  // row j read by its new owner before it is overwritten
    if (tid < kend - k) A[(int64_t)(j + tid) * ld + j] = u[k + tid];
  }
  if (own) {
    C2 l;
    l.x = 0;
    l.y = 0;
#pragma unroll
    for (int s = 0; s < IB; s++)
      if (s == k) l = v[s];
    const C2 piv = u[k];
    if (piv.x != 0 || piv.y != 0) {
      // l / piv (Smith's formula)
      const T a = piv.x, bb = piv.y;
      C2 q;
      if (fabs(a) >= fabs(bb)) {
        const T r = bb / a, den = a + bb * r;
        q.x = (l.x + l.y * r) / den;
        q.y = (l.y - l.x * r) / den;
      } else {
        const T r = a / bb, den = a * r + bb;
        q.x = (l.x * r + l.y) / den;
        q.y = (l.y * r - l.x) / den;
      }
      l = q;
    }
#pragma unroll
    for (int s = 0; s < IB; s++) {
      if (s == k) v[s] = l;
      if (s > k && s < kend) {
        const C2 uu = u[s];
        v[s].x -= l.x * uu.x - l.y * uu.y;
        v[s].y -= l.x * uu.y + l.y * uu.x;
      }
      if (s >= k && s < kend) A[(int64_t)(i0 + s) * ld + i] = v[s];
    }
  }
  if (!fuse_next) return;
  double cv = -1.0;
  int ci = n, cw = tid;
  if (own) {
#pragma unroll
    for (int s = 0; s < IB; s++)
      if (s == k + 1) cv = (double)cabs1<T>(v[s]);
    ci = (int)i;
  }
  block_argmax(cv, ci, cw, sv, si, sw);
  if (tid == cw) {
    const size_t slot = (size_t)(par ^ 1) * w.mb + b;
    w.val[slot] = cv;
    w.idx[slot] = ci;
  }
  if (tid == cw && ci < n) {
    const size_t slot = (size_t)(par ^ 1) * w.mb + b;
    C2 *row = (C2 *)w.row + slot * IB;
#pragma unroll
    for (int s = 0; s < IB; s++)
      if (s > k && s < kend) row[s] = v[s];
  }
}

// ---------------------------------------------------------------------------
// Persistent factorisation of one inner block (ib <= IB columns starting at
// i0) in a single cooperative launch.  Workgroup g owns the 256 row
// positions i0 + 256 g .. and keeps their ib values in registers for the
// whole block.  Per column the workgroups exchange data through tagged
// 64-bit words, (round << 32) | 32-bit payload, written and polled with
// relaxed device-scope atomics: a reader accepts a word only when it
// carries the current round, so no fences are needed, and the two slot
// parities keep round r+2 from overwriting data of round r before every
// workgroup has finished with it.  Per column j:
//   1. every workgroup publishes its candidate (largest |re|+|im| over its
//      positions >= j, lowest row on ties) with that row's ib values, and
//      the owner of position j publishes row j;
//   2. every workgroup reads all candidates and picks the same pivot p;
//   3. it reads the pivot row (and, in p's owner, old row j), swaps in
//      registers and applies the scaling and rank-1 update to its rows > j.
// Every wait is bounded: a poll that exceeds kSpinLimit records a timeout
// in sync.err and the workgroup leaves together (checked by the host).
typedef unsigned long long u64;
constexpr int kRowWords = 4 * IB;                 // ib complex128 values as 32-bit words
constexpr int kCandWords = 3 + kRowWords;         // magnitude (1 or 2 words), row index, row values
constexpr long kSpinLimit = 1L << 22;
constexpr int kPanelMaxGroups = 256;              // candidate slots (rows <= 256 * 256 per launch)

struct PanelSync {
  u64 *cand;  // [2][gmax][kCandWords]
  u64 *diag;  // [2][kRowWords]
  int gmax;
  int *err;
};

__device__ __forceinline__ void put_word(u64 *w, unsigned tag, unsigned payload) {
  __hip_atomic_store(w, ((u64)tag << 32) | payload, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Polls up to NW words at once (null pointers are skipped) until each
// carries `tag`: the loads of one pass are issued back to back, so a pass
// costs one memory round trip.  False on timeout (after recording it).
template <int NW>
__device__ __forceinline__ bool poll_words(u64 *const (&w)[NW], unsigned tag, unsigned (&out)[NW], int *err) {
  bool got[NW];
#pragma unroll
  for (int q = 0; q < NW; q++) got[q] = (w[q] == nullptr);
  for (long spins = 0;; spins++) {
    u64 x[NW];
#pragma unroll
    for (int q = 0; q < NW; q++)
      if (!got[q]) x[q] = __hip_atomic_load(w[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool all = true;
#pragma unroll
    for (int q = 0; q < NW; q++) {
      if (!got[q] && (unsigned)(x[q] >> 32) == tag) {
        out[q] = (unsigned)x[q];
        got[q] = true;
      }
      all = all && got[q];
    }
    if (all) return true;
    if ((spins & 255) == 255 &&
        (spins > kSpinLimit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      atomicExch(err, 1);
      return false;
    }
  }
}

// 32-bit words of one complex value of precision T
template <typename T> struct Words;
template <> struct Words<float> {
  static constexpr int N = 2, V = 1;  // words per complex value, per pivot magnitude
  static __device__ __forceinline__ unsigned mag(double m, int q) { return __float_as_uint((float)m); }
  static __device__ __forceinline__ double mag_of(const unsigned *w) { return (double)__uint_as_float(w[0]); }
  static __device__ __forceinline__ unsigned get(float2 v, int q) { return __float_as_uint(q ? v.y : v.x); }
  static __device__ __forceinline__ float2 make(const unsigned *w) {
    return make_float2(__uint_as_float(w[0]), __uint_as_float(w[1]));
  }
};
template <> struct Words<double> {
  static constexpr int N = 4, V = 2;
  static __device__ __forceinline__ unsigned mag(double m, int q) {
    return q ? (unsigned)__double2hiint(m) : (unsigned)__double2loint(m);
  }
  static __device__ __forceinline__ double mag_of(const unsigned *w) { return __hiloint2double((int)w[1], (int)w[0]); }
  static __device__ __forceinline__ unsigned get(double2 v, int q) {
    const double d = (q >> 1) ? v.y : v.x;
    return (q & 1) ? (unsigned)__double2hiint(d) : (unsigned)__double2loint(d);
  }
  static __device__ __forceinline__ double2 make(const unsigned *w) {
    return make_double2(__hiloint2double((int)w[1], (int)w[0]), __hiloint2double((int)w[3], (int)w[2]));
  }
};

// A workgroup that leaves on a timeout: workgroup 0 records identity
// interchanges for the columns [j, jend) it did not finish, so that the
// kernels after the panel only ever see pivot rows inside the matrix (the
// host reports the timeout after the factorisation).
__device__ __forceinline__ void leave_identity(int *ipiv, int g, int tid, int j, int jend) {
  if (g == 0 && tid == 0)
    for (int q = j; q < jend; q++) ipiv[q] = q + 1;
}

template <typename T>
__global__ __launch_bounds__(256) void panel_block_kernel(typename Cx<T>::type *A, int64_t ld, int i0, int ib,
                                                          int n, int *ipiv, int *info, PanelSync sync,
                                                          unsigned tag0) {
  using C2 = typename Cx<T>::type;
  constexpr int WN = Words<T>::N, VW = Words<T>::V;
  __shared__ unsigned uw[IB * WN], djw[IB * WN];    // pivot row / old row j as 32-bit words
  __shared__ unsigned pubw[IB * WN], pdw[IB * WN];  // rows this workgroup publishes
  __shared__ double sv[4];
  __shared__ int si[4], sw[4];
  __shared__ int fail;
  const int tid = threadIdx.x, g = blockIdx.x, G = gridDim.x;
  const int r = i0 + 256 * g + tid;
  const bool own = r < n;
  C2 v[IB];
#pragma unroll
  for (int s = 0; s < IB; s++) {
    v[s].x = 0;
    v[s].y = 0;
    if (own && s < ib) v[s] = A[(int64_t)(i0 + s) * ld + r];
  }
  if (tid == 0) fail = 0;
  __syncthreads();
  for (int k = 0; k < ib; k++) {
    const int j = i0 + k;
    const unsigned tag = tag0 + (unsigned)k;
    const int par = tag & 1;
    u64 *cand = sync.cand + (size_t)par * sync.gmax * kCandWords;
    u64 *diag = sync.diag + (size_t)par * kRowWords;
    // 1. local candidate and publication.  The magnitude of a complex64
    //    entry is exact in float, so that path sends one word for it.
    double val = -1.0;
    int idx = n, win = tid;
    if (own && r >= j) {
#pragma unroll
      for (int s = 0; s < IB; s++)
        if (s == k) val = (double)cabs1<T>(v[s]);
      idx = r;
    }
    block_argmax(val, idx, win, sv, si, sw);
    const int nw = WN * ib;
    u64 *slot = cand + (size_t)g * kCandWords;
    // value and row index every round (a workgroup without rows >= j
    // offers -1, which never wins)
    if (tid < VW) put_word(slot + tid, tag, Words<T>::mag(val, tid));
    if (tid == VW) put_word(slot + VW, tag, (unsigned)idx);
    // rows to publish go through LDS within the owning wave, so that its
    // lanes store one or two words each without a workgroup barrier
    if (idx < n && (tid >> 6) == (win >> 6)) {
      if (tid == win) {
#pragma unroll
        for (int s = 0; s < IB; s++)
          if (s < ib)
#pragma unroll
            for (int q = 0; q < WN; q++) pubw[WN * s + q] = Words<T>::get(v[s], q);
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (int w = tid & 63; w < nw; w += 64) put_word(slot + VW + 1 + w, tag, pubw[w]);
    }
    if (g == 0 && (tid >> 6) == ((j - i0) >> 6)) {  // position j lies in workgroup 0
      if (r == j) {
#pragma unroll
        for (int s = 0; s < IB; s++)
          if (s < ib)
#pragma unroll
            for (int q = 0; q < WN; q++) pdw[WN * s + q] = Words<T>::get(v[s], q);
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (int w = tid & 63; w < nw; w += 64) put_word(diag + w, tag, pdw[w]);
    }
    // 2. every workgroup reduces all candidates to the same pivot
    double bv = -1.0;
    int bi = n, bw = 0;
    if (tid < G) {
      u64 *slot = cand + (size_t)tid * kCandWords;
      u64 *const ws[3] = {slot, VW > 1 ? slot + 1 : nullptr, slot + VW};
      unsigned got[3];
      if (!poll_words<3>(ws, tag, got, sync.err)) {
        fail = 1;
      } else {
        const unsigned mw[2] = {got[0], got[1]};
        bv = Words<T>::mag_of(mw);
        bi = (int)got[2];
        bw = tid;
      }
    }
    block_argmax(bv, bi, bw, sv, si, sw);  // contains barriers: `fail` is visible after it
    if (fail) {
      leave_identity(ipiv, g, tid, j, i0 + ib);
      return;
    }
    // no candidate with a magnitude that compares (all rows NaN): keep row j,
    // whose values workgroup 0 publishes every round
    const bool none = !(bi >= j && bi < n);
    const int p = none ? j : bi;
    const bool zero = !(bv > 0.0);
    // 3. pivot row (and old row j for p's owner) into LDS: whole rows, the
    //    multipliers of columns i0..j-1 move with the interchange as well
    if (tid < nw) {
      const bool need_dj = p != j && (p - i0) / 256 == g;
      u64 *const ws[2] = {none ? diag + tid : cand + (size_t)bw * kCandWords + VW + 1 + tid,
                          need_dj ? diag + tid : nullptr};
      unsigned got[2];
      if (!poll_words<2>(ws, tag, got, sync.err)) {
        fail = 1;
      } else {
        uw[tid] = got[0];
        if (need_dj) djw[tid] = got[1];
      }
    }
    __syncthreads();
    if (fail) {
      leave_identity(ipiv, g, tid, j, i0 + ib);
      return;
    }
    if (g == 0 && tid == 0) {
      ipiv[j] = p + 1;
      if (zero) atomicCAS(info, 0, j + 1);
    }
    if (own && r == j) {
#pragma unroll
      for (int s = 0; s < IB; s++)
        if (s < ib) v[s] = Words<T>::make(uw + WN * s);
    } else if (own && p != j && r == p) {
#pragma unroll
      for (int s = 0; s < IB; s++)
        if (s < ib) v[s] = Words<T>::make(djw + WN * s);
    }
    if (own && r > j) {
      C2 l;
      l.x = 0;
      l.y = 0;
#pragma unroll
      for (int s = 0; s < IB; s++)
        if (s == k) l = v[s];
      const C2 piv = Words<T>::make(uw + WN * k);
      if (piv.x != 0 || piv.y != 0) {
        const T a = piv.x, bb = piv.y;
        C2 q;
        if (fabs(a) >= fabs(bb)) {
          const T rr = bb / a, den = a + bb * rr;
          q.x = (l.x + l.y * rr) / den;
          q.y = (l.y - l.x * rr) / den;
        } else {
          const T rr = a / bb, den = a * rr + bb;
          q.x = (l.x * rr + l.y) / den;
          q.y = (l.y * rr - l.x) / den;
        }
        l = q;
      }
#pragma unroll
      for (int s = 0; s < IB; s++) {
        if (s == k) v[s] = l;
        if (s > k && s < ib) {
          const C2 uu = Words<T>::make(uw + WN * s);
          v[s].x -= l.x * uu.x - l.y * uu.y;
          v[s].y -= l.x * uu.y + l.y * uu.x;
        }
      }
    }
    // uw / djw are next rewritten after the two barriers of the next
    // round's candidate reduction, which order these reads before that
  }
#pragma unroll
  for (int s = 0; s < IB; s++)
    if (own && s < ib) A[(int64_t)(i0 + s) * ld + r] = v[s];
}

// Row interchanges k1..k2-1 (ipiv 1-based, k2 - k1 <= NB) as one gather.
// The plan kernel replays the interchanges on slot indices: slot s < jb is
// row k1 + s, slot jb + k is row p_k for an interchange k whose pivot row
// lies below k2 (one slot per distinct such row, at its first occurrence).
// After the replay slot s holds the old content of slot idx[s]; the plan is
// the list of (destination row, source row) pairs, -1 for unchanged slots:
// plan[0] = 2 jb, plan[1 + e] = destination, plan[1 + 2 NB + e] = source.
// The plan of interchanges k1..k2-1 by the threads of one workgroup (at
// least NB of them; pv[NB], first[NB], idx[2 NB] in LDS).
__device__ void laswp_plan_block(const int *ipiv, int k1, int k2, int n, int *plan, int *pv, int *first, int *idx) {
  const int tid = threadIdx.x, jb = k2 - k1;
  if (tid < jb) {
    // a pivot row of column k lies in [k, n); anything else (a factorisation
    // left early) is replaced by no interchange
    const int p = ipiv[k1 + tid] - 1;
    pv[tid] = (p >= k1 + tid && p < n) ? p : k1 + tid;
  }
  const int nt = blockDim.x;
  for (int q = tid; q < 2 * NB; q += nt) idx[q] = q;
  __syncthreads();
  if (tid < jb) {
    int f = tid;
    if (pv[tid] >= k2)
      for (int q = 0; q < tid; q++)
        if (pv[q] == pv[tid]) {
          f = q;
          break;
        }
    first[tid] = f;
  }
  __syncthreads();
  if (tid == 0) {
    for (int k = 0; k < jb; k++) {
      const int p = pv[k];
      if (p == k1 + k) continue;
      const int b = (p < k2) ? p - k1 : jb + first[k];
      const int t = idx[k];
      idx[k] = idx[b];
      idx[b] = t;
    }
    plan[0] = 2 * jb;
  }
  __syncthreads();
  for (int e = tid; e < 2 * jb; e += nt) {
    int dst = -1, src = -1;
    const bool live = e < jb || (first[e - jb] == e - jb && pv[e - jb] >= k2);
    if (live) {
      const int se = idx[e];
      dst = e < jb ? k1 + e : pv[e - jb];
      src = se < jb ? k1 + se : pv[se - jb];
      if (src == dst) dst = -1;
    }
    plan[1 + e] = dst;
    plan[1 + 2 * NB + e] = src;
  }
}

__global__ __launch_bounds__(256) void laswp_plan_kernel(const int *ipiv, int k1, int k2, int n, int *plan) {
  __shared__ int pv[NB], first[NB], idx[2 * NB];
  laswp_plan_block(ipiv, k1, k2, n, plan, pv, first, idx);
}

// Applies a laswp plan to columns [c0, c1) and [c2, c3): each block gathers
// the source rows of kLaswpCols columns into LDS, then scatters them to the
// destination rows (the barrier orders every read before any write).
constexpr int kLaswpCols = 4;
template <typename T>
__global__ __launch_bounds__(256) void laswp_apply_kernel(typename Cx<T>::type *A, int64_t ld, const int *plan,
                                                          int c0, int c1, int c2, int c3) {
  using C2 = typename Cx<T>::type;
  __shared__ C2 buf[kLaswpCols][2 * NB];
  const int tid = threadIdx.x;
  const int m = plan[0];
  int dst[2], src[2];
#pragma unroll
  for (int u = 0; u < 2; u++) {
    const int e = tid + 256 * u;
    dst[u] = e < m ? plan[1 + e] : -1;
    src[u] = e < m ? plan[1 + 2 * NB + e] : -1;
  }
  const int n1 = c1 - c0, ncols = n1 + (c3 - c2);
  for (int cb = blockIdx.x * kLaswpCols; cb < ncols; cb += gridDim.x * kLaswpCols) {
#pragma unroll
    for (int q = 0; q < kLaswpCols; q++) {
      const int t = cb + q;
      if (t >= ncols) break;
      const C2 *col = A + (int64_t)(t < n1 ? c0 + t : c2 + (t - n1)) * ld;
#pragma unroll
      for (int u = 0; u < 2; u++)
        if (dst[u] >= 0) buf[q][tid + 256 * u] = col[src[u]];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kLaswpCols; q++) {
      const int t = cb + q;
      if (t >= ncols) break;
      C2 *col = A + (int64_t)(t < n1 ? c0 + t : c2 + (t - n1)) * ld;
#pragma unroll
      for (int u = 0; u < 2; u++)
        if (dst[u] >= 0) col[dst[u]] = buf[q][tid + 256 * u];
    }
    __syncthreads();
  }
}

// X := inv(L) X for the unit lower w x w block (w <= 32) at (r0, r0),
// columns [c0, c0 + ncols) of X.  L is staged in LDS; each half-wave holds
// one column of X (lane i = row i), and step k broadcasts x_k with a
// shuffle and updates the rows below it.
template <typename T>
__global__ __launch_bounds__(256) void trsm_unit_lower_kernel(typename Cx<T>::type *A, int64_t ld, int64_t r0,
                                                               int w, int64_t c0, int64_t ncols) {
  using C2 = typename Cx<T>::type;
  __shared__ C2 Ls[IB][IB];  // Ls[k][i] = L(r0 + i, r0 + k), zero outside the w x w block
  for (int e = threadIdx.x; e < IB * IB; e += 256) {
    const int k = e / IB, i = e % IB;
    T re = 0, im = 0;
    if (i < w && k < w && i > k) {
      const C2 l = A[(r0 + k) * ld + r0 + i];
      re = l.x;
      im = l.y;
    }
    Ls[k][i].x = re;
    Ls[k][i].y = im;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, i = lane & 31, half = lane >> 5;
  const int64_t pairs = (ncols + 1) / 2;
  for (int64_t cp = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); cp < pairs; cp += (int64_t)gridDim.x * 4) {
    const int64_t c = 2 * cp + half;
    const bool live = c < ncols && i < w;
    C2 *x = A + (c0 + c) * ld + r0 + i;
    T vr = 0, vi = 0;
    if (live) {
      const C2 v = *x;
      vr = v.x;
      vi = v.y;
    }
    for (int k = 0; k < w - 1; k++) {
      const int src = (lane & 32) + k;
      const T xr = __shfl(vr, src), xi = __shfl(vi, src);
      const T lr = Ls[k][i].x, li = Ls[k][i].y;
      vr -= lr * xr - li * xi;
      vi -= lr * xi + li * xr;
    }
    if (live) {
      C2 v;
      v.x = vr;
      v.y = vi;
      *x = v;
    }
  }
}

static unsigned trsm_grid(int64_t ncols) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(((ncols + 1) / 2 + 3) / 4, 2048));
}

// X := inv(L) X with L the unit lower w x w block at (r0, r0), w <= NB,
// processed in IB-row slices: small triangle then MFMA update of the rows
// below within the block.
template <typename T>
static int trsm_block(const Lane &L, typename Cx<T>::type *A, int64_t ld, int64_t r0, int64_t w, int64_t c0,
                      int64_t ncols, PhaseClock *st) {
  if (ncols <= 0) return 0;
  hipStream_t s = L.s;
  for (int64_t b = 0; b < w; b += IB) {
    const int bw = (int)std::min<int64_t>(IB, w - b);
    {
      Timed t(st, s, kPhTrsm);
      hipLaunchKernelGGL(trsm_unit_lower_kernel<T>, dim3(trsm_grid(ncols)), dim3(256), 0, s, A, ld, r0 + b, bw, c0,
                         ncols);
      MOM_HIP(hipGetLastError());
      MOM_TRY(t.done());
    }
    const int64_t below = w - b - bw;
    if (below > 0) {
      int rc = gemm_sub<T>(L, A + (r0 + b) * ld + r0 + b + bw, ld, A + c0 * ld + r0 + b, ld,
                           A + c0 * ld + r0 + b + bw, ld, below, ncols, bw, st);
      if (rc) return rc;
    }
  }
  return 0;
}

// E = I - inv(L11) for the unit lower triangular NB x NB diagonal block
// L11 of a factored panel (complex64, column-major, ld), written
// column-major with leading dimension NB (zero on and above the diagonal).
// With it the U block of the panel's row block is one depth-NB product:
// U12 = inv(L11) A12 = A12 - E A12.  One workgroup per kInvCols columns:
// its block column X of inv(L11) is formed in LDS in FP64 by block forward
// substitution, rows in blocks of 32: first the product with the rows above
// (L staged through LDS 32 columns at a time), then the 32 x 32 unit
// triangle; E is rounded to complex64 once.  The LDS and register footprint
// fits in the slot one trailing-update workgroup leaves, so the kernel runs
// beside the update on the look-ahead stream.
constexpr int kInvCols = 8;

__global__ __launch_bounds__(256) void unit_lower_inverse_kernel(const float2 *__restrict__ L, int64_t ld,
                                                                 float2 *__restrict__ E) {
  __shared__ double2 X[NB][kInvCols + 1];
  __shared__ double2 Lc[32][33];
  const int tid = threadIdx.x, c0 = kInvCols * blockIdx.x;
  const int c = tid % kInvCols, g = tid / kInvCols;  // column of the workgroup, row of a 32-row block
  const int b0 = c0 / 32;                             // row block holding the diagonal of these columns
  for (int b = b0; b < NB / 32; b++) {
    const int r0 = 32 * b;
    // acc = sum_{k in [32 b0, r0)} L[r0 + g][k] X[k][c] (X is zero above the diagonal)
    double2 acc = make_double2(0.0, 0.0);
    for (int kb = 32 * b0; kb < r0; kb += 32) {
      for (int e = tid; e < 1024; e += 256) {
        const int rr = e & 31, kk = e >> 5;
        const float2 l = L[(int64_t)(kb + kk) * ld + r0 + rr];
        Lc[rr][kk] = make_double2(l.x, l.y);
      }
      __syncthreads();
#pragma unroll 8
      for (int kk = 0; kk < 32; kk++) {
        const double2 x = X[kb + kk][c], l = Lc[g][kk];
        acc.x += l.x * x.x - l.y * x.y;
        acc.y += l.x * x.y + l.y * x.x;
      }
      __syncthreads();
    }
    // right-hand side of the block: identity rows minus the product
    for (int e = tid; e < 1024; e += 256) {
      const int rr = e & 31, kk = e >> 5;
      const float2 l = L[(int64_t)(r0 + kk) * ld + r0 + rr];
      Lc[rr][kk] = make_double2(l.x, l.y);
    }
    X[r0 + g][c] = make_double2((r0 + g == c0 + c ? 1.0 : 0.0) - acc.x, -acc.y);
    __syncthreads();
    // unit lower triangle of the block, column by column
    for (int kk = 0; kk < 31; kk++) {
      const double2 xk = X[r0 + kk][c];
      if (g > kk) {
        const double2 l = Lc[g][kk];
        double2 &t = X[r0 + g][c];
        t.x -= l.x * xk.x - l.y * xk.y;
        t.y -= l.x * xk.y + l.y * xk.x;
      }
      __syncthreads();
    }
  }
  for (int e = tid; e < NB * kInvCols; e += 256) {
    const int r = e % NB, cc = e / NB;
    float2 v = make_float2(0.f, 0.f);
    if (r > c0 + cc) {
      v.x = (float)-X[r][cc].x;
      v.y = (float)-X[r][cc].y;
    }
    E[(int64_t)(c0 + cc) * NB + r] = v;
  }
}

// ---------------------------------------------------------------------------
// Tournament pivoting for the complex64 panels (communication-avoiding LU,
// Grigori, Demmel and Xiang, SIAM J. Matrix Anal. Appl. 32 (2011)): the
// pivot rows of an inner block of ib <= 32 columns are chosen by partial
// pivoting inside workgroups and a reduction tree over their choices, so
// that the block needs a few launches instead of one grid-wide exchange per
// column.  The rows are then moved to the top and the block is factored
// without further pivoting (L21 = A21 inv(U11)).  The pivots differ from
// LAPACK's but are as good in practice; the complex64 factors only
// precondition the complex128 refinement, whose result does not depend on
// them beyond rounding.  The complex128 path keeps LAPACK's partial pivoting.
// Off by default (a system chooses it with mom_system_set_pivoting, the
// test hook mom_test_tournament forces it): LAPACK's pivots stay the
// default.  On the 15360-unknown ellipsoid the factorisation takes 0.16 s
// with it against 0.23 s with the cooperative partial-pivoting panel; one
// level of 32 selection steps takes about 40 us (1.2 us per column step,
// latency of one wave per SIMD).  Its launch sequence depends only on n, so
// a system replays it from a captured HIP graph (factor_graph).
//
// Selection, one workgroup of 256 threads per group of kTsluRows = 256
// rows, one row per thread in registers: level 0 takes 256 consecutive rows
// of the block, a later level up to eight candidate sets of ib rows (their
// original values read again from A).  Partial pivoting over the group --
// pivot search by a workgroup argmax of |re| + |im| with ties to the lowest
// row as IZAMAX, pivot row broadcast through LDS -- writes the ib rows
// chosen, in pivot order, as the group's candidate set (-1 where the group
// had fewer rows).  A NaN magnitude counts as 0, so such rows only win
// columns without any other nonzero candidate.  (Groups of 1024 rows, two
// per thread, make the tree one level shallower but each column step costs
// more than 1.5 times as much: slower on the 15360-unknown ellipsoid.)
constexpr int kTsluThreads = 256;
constexpr int kTsluRpt = 1;                              // rows per thread
constexpr int kTsluRows = kTsluThreads * kTsluRpt;       // rows per workgroup
constexpr int kTsluSets = kTsluRows / IB;                // candidate sets merged per workgroup
constexpr int kTsluWaves = kTsluThreads / 64;

__device__ __forceinline__ float2 cdiv_f(float2 a, float2 b) {
  if (b.x == 0.f && b.y == 0.f) return a;
  const float den = b.x * b.x + b.y * b.y;
  return make_float2((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}

// LAPACK-style interchanges of the chosen rows sel[0..ib) into positions
// i0..i0+ib-1 (by one whole wave; lane s keeps slot s of the positions moved so far
// and their rows): ipiv[i0 + k] = 1 + the position row sel[k] occupies after
// the interchanges of steps < k.  A row below the block is still at its own
// position until it is chosen; rows of the block may have moved.
__device__ void tslu_pivots_wave(const int *__restrict__ sel, int i0, int ib, int *ipiv) {
  const int lane = threadIdx.x & 63;
  int mpos = -1, mrow = -1, np = 0;
  for (int k = 0; k < ib; k++) {
    const int r = sel[k];
    const unsigned long long br = __ballot(lane < np && mrow == r);
    const int q = br ? __shfl(mpos, __builtin_ffsll((long long)br) - 1) : r;
    if (lane == 0) ipiv[i0 + k] = q + 1;
    if (q == i0 + k) continue;  // already in place
    const int p = i0 + k;
    const unsigned long long bp = __ballot(lane < np && mpos == p);
    const int a = bp ? __shfl(mrow, __builtin_ffsll((long long)bp) - 1) : p;  // row now at position p
    const unsigned long long bq = __ballot(lane < np && mpos == q);
    int sp = bp ? __builtin_ffsll((long long)bp) - 1 : np++;
    if (lane == sp) {
      mpos = p;
      mrow = r;
    }
    int sq = bq ? __builtin_ffsll((long long)bq) - 1 : np++;
    if (lane == sq) {
      mpos = q;
      mrow = a;
    }
  }
}

// Wave maximum / minimum of an int through DPP (quad permutations, row
// rotations, row broadcasts; lane 63 ends with the result over all lanes),
// read back to every lane; lanes without a valid DPP source keep their own
// value.  Non-negative floats order as their bits do as ints, below which
// every negative float lies, so the maximum of magnitudes (or -1 for none)
// is taken on their bits.
template <bool MAX>
__device__ __forceinline__ int wave_reduce_i32(int x) {
#define MOM_DPP_STEP(c)                                              \
  {                                                                  \
    const int t = __builtin_amdgcn_update_dpp(x, x, c, 0xf, 0xf, false); \
    x = MAX ? max(x, t) : min(x, t);                                 \
  }
  MOM_DPP_STEP(0xb1)   // quad_perm [1, 0, 3, 2]
  MOM_DPP_STEP(0x4e)   // quad_perm [2, 3, 0, 1]
  MOM_DPP_STEP(0x124)  // row_ror:4
  MOM_DPP_STEP(0x128)  // row_ror:8
  MOM_DPP_STEP(0x142)  // row_bcast:15
  MOM_DPP_STEP(0x143)  // row_bcast:31
#undef MOM_DPP_STEP
  return __builtin_amdgcn_readlane(x, 63);
}

// One barrier per column.  Each wave finds its largest magnitude by a DPP
// float maximum and the lowest row holding it (the first lane of the ballot,
// or a DPP row minimum when several lanes tie), publishes (magnitude, row)
// and, from the lane holding it, that row before the barrier; after it every
// thread takes the best of the kTsluWaves entries (larger magnitude, then
// lower row).  The buffers alternate by column parity, so a wave that runs
// ahead writes the other parity while the rest still read.  The multipliers
// use one reciprocal of |pivot|^2 per column.  With lu_top (the last level,
// one group) the winner of step k also stores its row -- multipliers left
// of k, U entries from k on -- as row k of the block's L11 U11, and a zero
// pivot is recorded in info (first one).
__global__ __launch_bounds__(kTsluThreads) void tslu_select_kernel(const float2 *__restrict__ A, int64_t ld, int i0,
                                                                   int ib, int n, const int *__restrict__ in,
                                                                   int nsets, int ngroups, int *__restrict__ out,
                                                                   float2 *__restrict__ lu_top,
                                                                   int *__restrict__ info, int *__restrict__ ipiv,
                                                                   int *__restrict__ plan) {
  static_assert(kTsluWaves == 4, "the best of the waves' candidates is read as one float4 and one int4");
  __shared__ __align__(16) float pmag[2][kTsluWaves];
  __shared__ __align__(16) int prw[2][kTsluWaves];
  __shared__ float2 prow[2][kTsluWaves][IB];
  __shared__ int pv[NB], first[NB], idx[2 * NB];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int grp = blockIdx.x;
  int row[kTsluRpt];
  bool act[kTsluRpt];
  float2 v[kTsluRpt][IB];
#pragma unroll
  for (int h = 0; h < kTsluRpt; h++) {
    const int q = tid + kTsluThreads * h;  // row of the group
    row[h] = -1;
    if (!in) {
      const int r = i0 + kTsluRows * grp + q;
      if (r < n) row[h] = r;
    } else {
      const int set = grp * kTsluSets + q / IB;
      if (set < nsets && q % IB < ib) row[h] = in[set * IB + q % IB];
    }
    act[h] = row[h] >= 0;
#pragma unroll
    for (int s = 0; s < IB; s++)
      v[h][s] = (act[h] && s < ib) ? A[(int64_t)(i0 + s) * ld + row[h]] : make_float2(0.f, 0.f);
  }
  // Straight-line steps (fully unrolled, static register indices; k < ib is
  // uniform): every row is updated whatever its state, with multiplier 0
  // once it is chosen or if it holds no row.  Rows already chosen are never
  // read again, so whatever such an update leaves in them is harmless.
#pragma clang loop unroll(full)
  for (int k = 0; k < IB; k++) {
    if (k >= ib) continue;
    const int par = k & 1;
    // this lane's candidate: magnitude (-1 without an active row) and row
    float cm = -1.f;
    int cr = 0x7fffffff, ch = 0;
#pragma unroll
    for (int h = 0; h < kTsluRpt; h++) {
      float c = fabsf(v[h][k].x) + fabsf(v[h][k].y);
      if (!(c == c)) c = 0.f;
      if (act[h] && (c > cm || (c == cm && row[h] < cr))) {
        cm = c;
        cr = row[h];
        ch = h;
      }
    }
    const float wm = __int_as_float(wave_reduce_i32<true>(__float_as_int(cm)));
    const unsigned long long tie = __ballot(cm >= 0.f && cm == wm);
    int wr = 0x7fffffff;
    if (tie) {
      wr = __builtin_amdgcn_readlane(cr, __builtin_ffsll((long long)tie) - 1);
      if (__builtin_popcountll(tie) > 1) wr = wave_reduce_i32<false>((tie >> lane) & 1ull ? cr : 0x7fffffff);
    }
    // columns s >= ib hold zeros in every row and stay zero, so the loops
    // below need no bound on s other than the static k
    if (tie && cm == wm && cr == wr) {
#pragma unroll
      for (int s = 0; s < IB; s++)
        if (s >= k) {
          float2 x = v[0][s];
#pragma unroll
          for (int h = 1; h < kTsluRpt; h++)
            if (ch == h) x = v[h][s];
          prow[par][wv][s] = x;
        }
    }
    if (lane == 0) {
      pmag[par][wv] = tie ? wm : -1.f;
      prw[par][wv] = wr;
    }
    __syncthreads();
    // the four waves' entries in two LDS reads, the best by selects
    const float4 pm4 = *reinterpret_cast<const float4 *>(&pmag[par][0]);
    const int4 pr4 = *reinterpret_cast<const int4 *>(&prw[par][0]);
    float bm = pm4.x;
    int br = pr4.x, bw = 0;
    auto take = [&](float m, int r, int q) {
      const bool better = m > bm || (m == bm && r < br);
      bm = better ? m : bm;
      br = better ? r : br;
      bw = better ? q : bw;
    };
    take(pm4.y, pr4.y, 1);
    take(pm4.z, pr4.z, 2);
    take(pm4.w, pr4.w, 3);
    // bm < 0: no rows left in this group (uniform); every row then has
    // multiplier 0 below
    if (bm < 0.f && tid == 0) out[grp * IB + k] = -1;
#pragma unroll
    for (int h = 0; h < kTsluRpt; h++) {
      if (bm >= 0.f && act[h] && row[h] == br) {
        out[grp * IB + k] = br;
        act[h] = false;
        if (lu_top) {
#pragma unroll
          for (int s = 0; s < IB; s++) lu_top[k * IB + s] = v[h][s];
          if (v[h][k].x == 0.f && v[h][k].y == 0.f) atomicCAS(info, 0, i0 + k + 1);
        }
      }
    }
    const float2 p = prow[par][bw][k];
    const float den = p.x * p.x + p.y * p.y;
    // the hardware reciprocal (1 ulp): the factors only precondition the
    // complex128 refinement
    const float inv = den != 0.f ? __builtin_amdgcn_rcpf(den) : 0.f;
    float2 l[kTsluRpt];
#pragma unroll
    for (int h = 0; h < kTsluRpt; h++) {
      // a zero pivot leaves the column as it is (as LAPACK's update skips it)
      const float2 a = v[h][k];
      const float2 q = den != 0.f ? make_float2((a.x * p.x + a.y * p.y) * inv, (a.y * p.x - a.x * p.y) * inv) : a;
      l[h] = act[h] ? q : make_float2(0.f, 0.f);
      v[h][k] = act[h] ? q : a;
    }
    // v -= l u by four FMAs per entry
#pragma unroll
    for (int s = 0; s < IB; s++) {
      if (s <= k) continue;
      const float2 u = prow[par][bw][s];
#pragma unroll
      for (int h = 0; h < kTsluRpt; h++) {
        const float nlx = -l[h].x, nly = -l[h].y;
        v[h][s].x = fmaf(l[h].y, u.y, fmaf(nlx, u.x, v[h][s].x));
        v[h][s].y = fmaf(nly, u.x, fmaf(nlx, u.y, v[h][s].y));
      }
    }
  }
  if (lu_top) {
    // the last level also turns its choice into LAPACK-style interchanges
    // and their laswp plan (out and then ipiv written by this workgroup)
    __syncthreads();
    if (wv == 0) tslu_pivots_wave(out, i0, ib, ipiv);
    __syncthreads();
    laswp_plan_block(ipiv, i0, i0 + ib, n, plan, pv, first, idx);
  }
}

// The inner block after its interchanges: the last selection level has
// factored the chosen rows in pivot order (lu_top, ib x ib, L11 U11 by
// rows); one launch stores it over the top block (workgroup 0) and solves
// l U11 = a_row in place for every row below the block, one per thread,
// with U11 from lu_top in LDS.  A zero pivot is skipped, as the partial
// pivoting panel does.
__global__ __launch_bounds__(256) void tslu_below_kernel(float2 *A, int64_t ld, int i0, int ib, int n,
                                                         const float2 *__restrict__ lu_top) {
  __shared__ float2 U[IB][IB + 1];
  const int tid = threadIdx.x;
  for (int e = tid; e < IB * IB; e += 256) {
    const int r = e / IB, c = e % IB;
    const float2 x = (r < ib && c < ib) ? lu_top[r * IB + c] : make_float2(0.f, 0.f);
    U[r][c] = r <= c ? x : make_float2(0.f, 0.f);
    if (blockIdx.x == 0 && r < ib && c < ib) A[(int64_t)(i0 + c) * ld + i0 + r] = x;
  }
  __syncthreads();
  const int64_t row = (int64_t)i0 + ib + (int64_t)blockIdx.x * 256 + tid;
  if (row >= n) return;
  float2 a[IB];
#pragma unroll
  for (int s = 0; s < IB; s++) a[s] = s < ib ? A[(int64_t)(i0 + s) * ld + row] : make_float2(0.f, 0.f);
#pragma unroll
  for (int s = 0; s < IB; s++) {
    float2 x = a[s];
#pragma unroll
    for (int t = 0; t < IB; t++) {
      if (t >= s) continue;
      const float2 u = U[t][s];
      x.x -= a[t].x * u.x - a[t].y * u.y;
      x.y -= a[t].x * u.y + a[t].y * u.x;
    }
    a[s] = cdiv_f(x, U[s][s]);  // columns s >= ib: zero over a zero pivot, kept zero
  }
#pragma unroll
  for (int s = 0; s < IB; s++)
    if (s < ib) A[(int64_t)(i0 + s) * ld + row] = a[s];
}

// the test hook's process-wide switch to tournament pivoting for complex64
// panels (a system chooses its own with mom_system_set_pivoting)
static int g_tournament = 0;

// Cooperative inner-block panel factorisation: synchronisation words, the
// largest co-resident grid of panel_block_kernel<T> (0: not available) and
// the next free round tag.
struct PanelCoop {
  PanelSync sync;
  int gmax = 0;
  unsigned *tag = nullptr;
  // set when the runtime refuses a cooperative launch (another process
  // holds the device's cooperative resources): the panels then take the
  // one-launch-per-column path
  bool *refused = nullptr;
};

// Factorisation of the panel of block columns [j0, j0+jb) over rows
// [j0, n) on lane L: inner blocks of IB columns, each factored (one
// cooperative launch, or one launch per column), its row interchanges
// carried across the other panel columns, then the U block and update of
// the panel's remaining columns.
template <typename T>
static int factor_panel(const Lane &L, typename Cx<T>::type *A, int64_t ld, int64_t n, int64_t j0, int64_t jb,
                        int *ipiv, PivotWork w, PhaseClock *st, const PanelCoop &coop) {
  hipStream_t s = L.s;
  for (int64_t i0 = j0; i0 < j0 + jb; i0 += IB) {
    const int64_t ib = std::min<int64_t>(IB, j0 + jb - i0);
    Timed tp(st, s, kPhPanel);
    const int64_t G = (n - i0 + 255) / 256;
    bool coop_done = false, tournament = false;
    if (std::is_same<T, float>::value && w.tournament) {
      // tournament over the rows, interchanges on every panel column, then
      // the block without pivoting
      int *cur = w.tsel, *nxt = w.tsel + (size_t)w.mb * IB;
      int64_t groups = (n - i0 + kTsluRows - 1) / kTsluRows;
      if constexpr (std::is_same<T, float>::value) {
        // the last level (one group) also factors the chosen rows into lu_top
        hipLaunchKernelGGL(tslu_select_kernel, dim3((unsigned)groups), dim3(kTsluThreads), 0, s, A, ld, (int)i0,
                           (int)ib, (int)n, (const int *)nullptr, 0, (int)groups, cur,
                           groups == 1 ? w.lu_top : (float2 *)nullptr, w.info, ipiv, w.plan);
        while (groups > 1) {
          const int64_t gn = (groups + kTsluSets - 1) / kTsluSets;
          hipLaunchKernelGGL(tslu_select_kernel, dim3((unsigned)gn), dim3(kTsluThreads), 0, s, A, ld, (int)i0,
                             (int)ib, (int)n, (const int *)cur, (int)groups, (int)gn, nxt,
                             gn == 1 ? w.lu_top : (float2 *)nullptr, w.info, ipiv, w.plan);
          std::swap(cur, nxt);
          groups = gn;
        }
        const int64_t lg = (jb + kLaswpCols - 1) / kLaswpCols;
        hipLaunchKernelGGL(laswp_apply_kernel<T>, dim3((unsigned)lg), dim3(256), 0, s, A, ld, w.plan, (int)j0,
                           (int)(j0 + jb), 0, 0);
        const int64_t below = n - i0 - ib;
        hipLaunchKernelGGL(tslu_below_kernel, dim3((unsigned)std::max<int64_t>(1, (below + 255) / 256)), dim3(256),
                           0, s, A, ld, (int)i0, (int)ib, (int)n, w.lu_top);
      }
      MOM_HIP(hipGetLastError());
      tournament = true;
    } else if (G <= coop.gmax && !(coop.refused && *coop.refused)) {
      // the whole inner block in one cooperative launch, then its row
      // interchanges on the other columns of the panel
      unsigned tag0 = *coop.tag;
      *coop.tag += 64;
      int i0i = (int)i0, ibi = (int)ib, ni = (int)n;
      int *infop = w.info;
      typename Cx<T>::type *Ap = A;
      int64_t ldv = ld;
      int *ipivp = ipiv;
      PanelSync sy = coop.sync;
      void *args[] = {&Ap, &ldv, &i0i, &ibi, &ni, &ipivp, &infop, &sy, &tag0};
      const hipError_t e = hipLaunchCooperativeKernel((const void *)panel_block_kernel<T>, dim3((unsigned)G),
                                                      dim3(256), args, 0, s);
      if (e == hipSuccess) {
        coop_done = true;
      } else if (coop.refused) {
        (void)hipGetLastError();
        *coop.refused = true;
      } else {
        return fail(MOM_ERR_HIP, std::string("hipLaunchCooperativeKernel(panel): ") + hipGetErrorString(e));
      }
    }
    if (tournament) {
      // interchanges and factors done above
    } else if (coop_done) {
      const int64_t left = i0 - j0, right = j0 + jb - (i0 + ib);
      if (left + right > 0) {
        hipLaunchKernelGGL(laswp_plan_kernel, dim3(1), dim3(256), 0, s, ipiv, (int)i0, (int)(i0 + ib), (int)n, w.plan);
        MOM_HIP(hipGetLastError());
        const int64_t groups = (left + right + kLaswpCols - 1) / kLaswpCols;
        hipLaunchKernelGGL(laswp_apply_kernel<T>, dim3((unsigned)std::min<int64_t>(groups, 4096)), dim3(256), 0, s,
                           A, ld, w.plan, (int)j0, (int)i0, (int)(i0 + ib), (int)(j0 + jb));
        MOM_HIP(hipGetLastError());
      }
    } else {
      // candidates of the block's first column, then one launch per column
      hipLaunchKernelGGL(pivot_candidates_kernel<T>, dim3((unsigned)((n - i0 + 255) / 256)), dim3(256), 0, s, A, ld,
                         (int)i0, (int)n, (int)i0, (int)(i0 + ib), w, (int)(i0 & 1));
      MOM_HIP(hipGetLastError());
      int nb_in = (int)((n - i0 + 255) / 256);
      for (int64_t j = i0; j < i0 + ib; j++) {
        const int64_t rows = n - j - 1;
        const int nblk = (int)std::max<int64_t>(1, (rows + 255) / 256);
        const int fuse = (j + 1 < i0 + ib) ? 1 : 0;
        hipLaunchKernelGGL(panel_column_kernel<T>, dim3((unsigned)nblk), dim3(256), 0, s, A, ld, (int)j, (int)n,
                           (int)i0, (int)(i0 + ib), (int)j0, (int)(j0 + jb), ipiv, w, nb_in, fuse);
        MOM_HIP(hipGetLastError());
        nb_in = nblk;
      }
    }
    MOM_TRY(tp.done());
    const int64_t rest = j0 + jb - (i0 + ib);
    if (rest > 0) {
      // U block of the remaining panel columns, then their update
      Timed tt(st, s, kPhTrsm);
      hipLaunchKernelGGL(trsm_unit_lower_kernel<T>, dim3(trsm_grid(rest)), dim3(256), 0, s, A, ld, i0, (int)ib,
                         i0 + ib, rest);
      MOM_HIP(hipGetLastError());
      MOM_TRY(tt.done());
      int rc = gemm_sub<T>(L, A + i0 * ld + i0 + ib, ld, A + (i0 + ib) * ld + i0, ld, A + (i0 + ib) * ld + i0 + ib,
                           ld, n - i0 - ib, rest, ib, st);
      if (rc) return rc;
    }
  }
  return 0;
}

// Look-ahead of the blocked LU (on unless switched off by the test hook).
static int g_lookahead = 1;
// U blocks of the complex64 LU: 1 (default) blocked substitution, 0 one
// product with E = I - inv(L11) where the shapes allow it (test hook).  The
// product saves about a third of the U-block time but loses about a decade
// of the factors' backward error on the PMCHWT matrices (|E| |A12| is much
// larger than |U12| there), which costs more refinement than it saves.
static int g_u_subst = 1;

// LU with partial pivoting of the n x n matrix A (ld), in place, blocked
// right-looking with look-ahead of depth one.  Step k, once panel k is
// factored: its row interchanges on the columns left and right of it and the
// U12 row block on the main stream; then the trailing update splits in two.
// The next block column is updated on the look-ahead stream, which goes on
// to factor panel k+1 there, while the main stream updates the remaining
// columns; the main stream waits for panel k+1 before step k+1.  The
// look-ahead stream also forms panel k+1's interchange plan and applies it
// to the columns left of block k (the main stream's update reads block k's
// columns and writes those right of block k+1), so that the main stream's
// chain before its next update holds only the interchanges of block k's
// columns and of the right part.  The two streams touch disjoint columns
// between fork and join.
template <typename T>
static int getrf(mom_ctx *ctx, typename Cx<T>::type *A, int64_t ld, int64_t n, int *ipiv, PivotWork w,
                 PhaseClock *st, PanelCoop coop = PanelCoop()) {
  const Lane M = main_lane(ctx);
  hipStream_t s = M.s;
  Lane X = M;
  const bool la = g_lookahead != 0 && n > NB;
  if (la) MOM_TRY(aux_lane(ctx, &X));
  // complex64: U blocks by one product with E = I - inv(L11) (E is formed
  // after each panel on the panel's stream, read by the next U block)
  bool inv_u = false;
  if constexpr (std::is_same<T, float>::value) {
    if (n > NB) {
      if (ctx->ublk.bytes < sizeof(float2) * NB * NB) MOM_TRY(ctx->ublk.alloc(sizeof(float2) * NB * NB));
      inv_u = g_u_subst == 0 && dma_gemm_ok(NB, NB, NB, ld, ctx->ublk.ptr, A + NB * ld) &&
              (reinterpret_cast<uintptr_t>(A) & 15) == 0;
    }
  }
  auto panel = [&](const Lane &P, int64_t j0, int64_t jb) -> int {
    MOM_TRY(factor_panel<T>(P, A, ld, n, j0, jb, ipiv, w, st, coop));
    if constexpr (std::is_same<T, float>::value) {
      if (inv_u && jb == NB && j0 + jb < n) {
        Timed tt(st, P.s, kPhTrsm);
        hipLaunchKernelGGL(unit_lower_inverse_kernel, dim3(NB / kInvCols), dim3(256), 0, P.s, A + j0 * ld + j0, ld,
                           ctx->ublk.as<float2>());
        MOM_HIP(hipGetLastError());
        MOM_TRY(tt.done());
      }
    }
    return 0;
  };
  auto laswp = [&](hipStream_t ls, int64_t c0, int64_t c1, int64_t c2, int64_t c3) -> int {
    const int64_t cols = (c1 - c0) + (c3 - c2);
    if (cols <= 0) return 0;
    const int64_t groups = (cols + kLaswpCols - 1) / kLaswpCols;
    hipLaunchKernelGGL(laswp_apply_kernel<T>, dim3((unsigned)std::min<int64_t>(groups, 4096)), dim3(256), 0, ls, A, ld,
                       w.plan, (int)c0, (int)c1, (int)c2, (int)c3);
    MOM_HIP(hipGetLastError());
    return 0;
  };
  if (la) {
    const int64_t jb0 = std::min<int64_t>(NB, n);
    MOM_TRY(panel(M, 0, jb0));
    hipLaunchKernelGGL(laswp_plan_kernel, dim3(1), dim3(256), 0, s, ipiv, 0, (int)jb0, (int)n, w.plan);
    MOM_HIP(hipGetLastError());
  }
  for (int64_t j0 = 0, prev = 0; j0 < n; prev = std::min<int64_t>(NB, n - j0), j0 += NB) {
    const int64_t jb = std::min<int64_t>(NB, n - j0);
    if (!la) MOM_TRY(panel(M, j0, jb));
    // row interchanges of this panel on the columns left and right of it
    // (with look-ahead the plan was formed, and the columns left of the
    // previous block interchanged, on the look-ahead stream)
    {
      Timed tl(st, s, kPhLaswp);
      if (la) {
        MOM_TRY(laswp(s, j0 - prev, j0, j0 + jb, n));
      } else if (j0 + (n - (j0 + jb)) > 0) {
        hipLaunchKernelGGL(laswp_plan_kernel, dim3(1), dim3(256), 0, s, ipiv, (int)j0, (int)(j0 + jb), (int)n,
                           w.plan);
        MOM_HIP(hipGetLastError());
        MOM_TRY(laswp(s, 0, j0, j0 + jb, n));
      }
      MOM_TRY(tl.done());
    }
    const int64_t right = n - (j0 + jb);
    if (right <= 0) continue;
    if (inv_u && jb == NB) {
      typename Cx<T>::type *A12 = A + (j0 + jb) * ld + j0;
      MOM_TRY(gemm_sub<T>(M, reinterpret_cast<const typename Cx<T>::type *>(ctx->ublk.ptr), NB, A12, ld, A12, ld, NB,
                          right, NB, st, true));
    } else {
      MOM_TRY(trsm_block<T>(M, A, ld, j0, jb, j0 + jb, right, st));
    }
    const int64_t r0 = j0 + jb, rows = n - r0;
    const typename Cx<T>::type *L21 = A + j0 * ld + r0, *U12 = A + r0 * ld + j0;
    typename Cx<T>::type *A22 = A + r0 * ld + r0;
    if (!la) {
      MOM_TRY(gemm_sub<T>(M, L21, ld, U12, ld, A22, ld, rows, right, jb, st));
      continue;
    }
    const int64_t nxt = std::min<int64_t>(NB, right);
    MOM_HIP(hipEventRecord(ctx->fork_ev, s));
    MOM_HIP(hipStreamWaitEvent(X.s, ctx->fork_ev, 0));
    MOM_TRY(gemm_sub<T>(X, L21, ld, U12, ld, A22, ld, rows, nxt, jb, st));
    MOM_TRY(panel(X, r0, nxt));
    // the next panel's plan, and its interchanges on the columns left of
    // this block (the main stream reads this block's columns until the join).
    // The join precedes those interchanges: the main stream's next step
    // touches only the columns from this block on, so it waits for the
    // panel and the plan alone.
    hipLaunchKernelGGL(laswp_plan_kernel, dim3(1), dim3(256), 0, X.s, ipiv, (int)r0, (int)(r0 + nxt), (int)n, w.plan);
    MOM_HIP(hipGetLastError());
    MOM_HIP(hipEventRecord(ctx->join_ev, X.s));
    MOM_TRY(laswp(X.s, 0, j0, j0, j0));
    MOM_TRY(gemm_sub<T>(M, L21, ld, U12 + nxt * ld, ld, A22 + nxt * ld, ld, rows, right - nxt, jb, st));
    MOM_HIP(hipStreamWaitEvent(s, ctx->join_ev, 0));
  }
  if (la) {
    // the last interchanges on the left columns
    MOM_HIP(hipEventRecord(ctx->join_ev, X.s));
    MOM_HIP(hipStreamWaitEvent(s, ctx->join_ev, 0));
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Inverses of the NB x NB diagonal blocks of L (unit lower) and U (upper),
// stored as nblk consecutive NB x NB column-major blocks in complex128.
// Grid (nblk, NB / kInvCols, 2): one workgroup per block, kInvCols columns
// of the inverse and factor (z = 0: L, 1: U).  The block column X is formed
// in LDS in FP64 by block substitution over rows in blocks of 32 (forward
// for L, backward for U): the product with the rows already solved (the
// factor staged through LDS 32 columns at a time), then the 32 x 32
// triangle.  Outside the matrix (a last block narrower than NB) the factor
// reads as the identity, so those columns of the inverse are unit vectors.
template <typename T>
__global__ __launch_bounds__(256) void diag_inverse_kernel(const typename Cx<T>::type *__restrict__ A, int64_t ld,
                                                            int64_t n, double2 *Linv, double2 *Uinv) {
  __shared__ double2 X[NB][kInvCols + 1];
  __shared__ double2 Fc[32][33];
  const int blk = blockIdx.x, c0 = kInvCols * blockIdx.y;
  const bool upper = blockIdx.z == 1;
  const int64_t d0 = (int64_t)blk * NB;
  const int w = (int)std::min<int64_t>(NB, n - d0);
  const int tid = threadIdx.x, c = tid % kInvCols, g = tid / kInvCols;
  // factor entry (r, k) of the diagonal block, identity outside the matrix
  auto F = [&](int r, int k) -> double2 {
    if (r >= w || k >= w) return make_double2(r == k ? 1.0 : 0.0, 0.0);
    if (!upper && r == k) return make_double2(1.0, 0.0);
    const typename Cx<T>::type v = A[(d0 + k) * ld + d0 + r];
    return make_double2(v.x, v.y);
  };
  const int bd = c0 / 32;                          // row block holding the diagonal of these columns
  const int bfirst = upper ? (c0 + kInvCols - 1) / 32 : bd;
  const int bstep = upper ? -1 : 1;
  const int blast = upper ? -1 : NB / 32;
  for (int b = bfirst; b != blast; b += bstep) {
    const int r0 = 32 * b;
    double2 acc = make_double2(0.0, 0.0);
    // rows already solved: L, blocks [bd, b); U, blocks (b, bfirst]
    const int klo = upper ? r0 + 32 : 32 * bd, khi = upper ? 32 * (bfirst + 1) : r0;
    for (int kb = klo; kb < khi; kb += 32) {
      for (int e = tid; e < 1024; e += 256) {
        const int rr = e & 31, kk = e >> 5;
        Fc[rr][kk] = F(r0 + rr, kb + kk);
      }
      __syncthreads();
#pragma unroll 8
      for (int kk = 0; kk < 32; kk++) {
        const double2 x = X[kb + kk][c], f = Fc[g][kk];
        acc.x += f.x * x.x - f.y * x.y;
        acc.y += f.x * x.y + f.y * x.x;
      }
      __syncthreads();
    }
    for (int e = tid; e < 1024; e += 256) {
      const int rr = e & 31, kk = e >> 5;
      Fc[rr][kk] = F(r0 + rr, r0 + kk);
    }
    X[r0 + g][c] = make_double2((r0 + g == c0 + c ? 1.0 : 0.0) - acc.x, -acc.y);
    __syncthreads();
    if (!upper) {
      for (int kk = 0; kk < 31; kk++) {
        const double2 xk = X[r0 + kk][c];
        if (g > kk) {
          const double2 f = Fc[g][kk];
          double2 &t = X[r0 + g][c];
          t.x -= f.x * xk.x - f.y * xk.y;
          t.y -= f.x * xk.y + f.y * xk.x;
        }
        __syncthreads();
      }
    } else {
      for (int kk = 31; kk >= 0; kk--) {
        if (g == kk) {  // divide row kk by the diagonal
          const double2 d = Fc[kk][kk], v = X[r0 + kk][c];
          const double den = d.x * d.x + d.y * d.y;
          X[r0 + kk][c] = make_double2((v.x * d.x + v.y * d.y) / den, (v.y * d.x - v.x * d.y) / den);
        }
        __syncthreads();
        const double2 xk = X[r0 + kk][c];
        if (g < kk) {
          const double2 f = Fc[g][kk];
          double2 &t = X[r0 + g][c];
          t.x -= f.x * xk.x - f.y * xk.y;
          t.y -= f.x * xk.y + f.y * xk.x;
        }
        __syncthreads();
      }
    }
  }
  double2 *out = (upper ? Uinv : Linv) + (int64_t)blk * NB * NB;
  for (int e = tid; e < NB * kInvCols; e += 256) {
    const int r = e % NB, cc = e / NB;
    const bool inside = upper ? r <= c0 + cc && r < 32 * (bfirst + 1) : r >= 32 * bd;
    out[(int64_t)(c0 + cc) * NB + r] = inside ? X[r][cc] : make_double2(0.0, 0.0);
  }
}

// Solves with the LU factors.  The factors are stored in TS (complex64 or
// complex128); vectors and arithmetic are in TC.  With TS = float, TC =
// double the preconditioner of the refinement is an exact FP64 linear
// operator built from the stored FP32 factors.
template <typename TS, typename TC>
__device__ __forceinline__ void cmac(typename Cx<TC>::type &acc, typename Cx<TS>::type m, typename Cx<TC>::type v) {
  const TC mr = (TC)m.x, mi = (TC)m.y;
  acc.x += mr * v.x - mi * v.y;
  acc.y += mr * v.y + mi * v.x;
}

// b[c0 + i, j] = xb[i, j] for i < w: a solved block back into its rows
template <typename CV>
__global__ void copy_block_kernel(const CV *__restrict__ xb, CV *b, int64_t ldb, int64_t c0, int w) {
  const int64_t rhs = blockIdx.y;
  for (int i = threadIdx.x; i < w; i += blockDim.x) b[rhs * ldb + c0 + i] = xb[rhs * NB + i];
}

// y[r] = sum_c M[r, c] x[c] for the NB x NB diagonal-block inverse M, into
// xb.  Grid (NB/8, nrhs): 8 rows per workgroup (128-byte row segments of M
// per column), 32 column slices per row, so that 32 workgroups share the
// 1 MB read of M.  Every workgroup reads the whole block x, so the result
// must not go back into x in this launch (a workgroup that starts late
// would read rows another one has already replaced): the next launch
// (panel_gemv_sub_kernel / _h_kernel, or copy_block_kernel after the last
// block) copies xb into the block's rows.
template <typename TS, typename TC, bool H = false>
__global__ __launch_bounds__(256) void block_apply_kernel(const double2 *__restrict__ M,
                                                           const typename Cx<TC>::type *__restrict__ x, int64_t ldx,
                                                           typename Cx<TC>::type *__restrict__ xb, int w) {
  using CV = typename Cx<TC>::type;
  __shared__ CV xs[NB];
  __shared__ CV part[32][8];
  const int64_t rhs = blockIdx.y;
  for (int i = threadIdx.x; i < NB; i += 256) xs[i] = (i < w) ? x[rhs * ldx + i] : CV{0, 0};
  __syncthreads();
  const int rl = threadIdx.x & 7, sl = threadIdx.x >> 3;
  const int r = blockIdx.x * 8 + rl;
  CV acc{0, 0};
  if (r < w) {
#pragma unroll 4
    for (int c = sl; c < w; c += 32) {
      // H: the conjugate transpose of M
      double2 m = H ? M[(int64_t)r * NB + c] : M[(int64_t)c * NB + r];
      if (H) m.y = -m.y;
      cmac<double, TC>(acc, m, xs[c]);
    }
  }
  part[sl][rl] = acc;
  __syncthreads();
  if (sl == 0 && r < w) {
    CV s = part[0][rl];
    for (int k = 1; k < 32; k++) {
      s.x += part[k][rl].x;
      s.y += part[k][rl].y;
    }
    xb[rhs * NB + r] = s;
  }
}

// b[r] -= sum_{c < w} A[r, c0 + c] xb[c] for rows [r0, r1); 64 rows x 4
// column slices per workgroup.
template <typename TS, typename TC>
__global__ __launch_bounds__(256) void panel_gemv_sub_kernel(const typename Cx<TS>::type *__restrict__ A, int64_t ld,
                                                              int64_t r0, int64_t r1, int64_t c0, int w,
                                                              const typename Cx<TC>::type *__restrict__ xb,
                                                              typename Cx<TC>::type *b, int64_t ldb) {
  using CV = typename Cx<TC>::type;
  __shared__ CV xs[NB];
  __shared__ CV part[4][64];
  const int64_t rhs = blockIdx.y;
  for (int i = threadIdx.x; i < w; i += 256) {
    const CV v = xb[rhs * NB + i];
    xs[i] = v;
    // the solved block into its own rows [c0, c0 + w), outside [r0, r1)
    if (blockIdx.x == 0) b[rhs * ldb + c0 + i] = v;
  }
  __syncthreads();
  const int rl = threadIdx.x & 63, sl = threadIdx.x >> 6;
  const int64_t r = r0 + blockIdx.x * 64 + rl;
  CV acc{0, 0};
  if (r < r1) {
    const typename Cx<TS>::type *a = A + c0 * ld + r;
#pragma unroll 4
    for (int c = sl; c < w; c += 4) cmac<TS, TC>(acc, a[(int64_t)c * ld], xs[c]);
  }
  part[sl][rl] = acc;
  __syncthreads();
  if (sl == 0 && r < r1) {
    CV s = part[0][rl];
    for (int k = 1; k < 4; k++) {
      s.x += part[k][rl].x;
      s.y += part[k][rl].y;
    }
    CV *p = b + rhs * ldb + r;
    CV o = *p;
    o.x -= s.x;
    o.y -= s.y;
    *p = o;
  }
}

template <typename TV>
__global__ void gather_rows_kernel(const TV *__restrict__ src, int64_t lds, TV *__restrict__ dst, int64_t ldd,
                                   const int *__restrict__ perm, int64_t n, int64_t nrhs) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * nrhs) return;
  const int64_t i = t % n, r = t / n;
  dst[r * ldd + i] = src[r * lds + perm[i]];
}

// Solve L U x = P b in place on B (n x nrhs, ldb, TC); perm is the row
// gather (B_perm[i] = B[perm[i]]); tmp is n x nrhs and xb NB x nrhs scratch.
template <typename TS, typename TC>
static int getrs(hipStream_t s, const typename Cx<TS>::type *LU, int64_t ld, int64_t n, const int *perm,
                 const double2 *Linv, const double2 *Uinv, typename Cx<TC>::type *B,
                 int64_t ldb, int64_t nrhs, typename Cx<TC>::type *tmp, typename Cx<TC>::type *xb) {
  using CV = typename Cx<TC>::type;
  hipLaunchKernelGGL(gather_rows_kernel<CV>, dim3((unsigned)((n * nrhs + 255) / 256)), dim3(256), 0, s, B, ldb, tmp,
                     n, perm, n, nrhs);
  MOM_HIP(hipGetLastError());
  const int64_t nblk = (n + NB - 1) / NB;
  for (int64_t b = 0; b < nblk; b++) {
    const int64_t r0 = b * NB;
    const int w = (int)std::min<int64_t>(NB, n - r0);
    hipLaunchKernelGGL((block_apply_kernel<TS, TC>), dim3(NB / 8, (unsigned)nrhs), dim3(256), 0, s,
                       Linv + b * NB * NB, tmp + r0, n, xb, w);
    if (r0 + w < n)
      hipLaunchKernelGGL((panel_gemv_sub_kernel<TS, TC>), dim3((unsigned)((n - r0 - w + 63) / 64), (unsigned)nrhs),
                         dim3(256), 0, s, LU, ld, r0 + w, n, r0, w, xb, tmp, n);
    else
      hipLaunchKernelGGL(copy_block_kernel<CV>, dim3(1, (unsigned)nrhs), dim3(256), 0, s, xb, tmp, n, r0, w);
  }
  for (int64_t b = nblk - 1; b >= 0; b--) {
    const int64_t r0 = b * NB;
    const int w = (int)std::min<int64_t>(NB, n - r0);
    hipLaunchKernelGGL((block_apply_kernel<TS, TC>), dim3(NB / 8, (unsigned)nrhs), dim3(256), 0, s,
                       Uinv + b * NB * NB, tmp + r0, n, xb, w);
    if (r0 > 0)
      hipLaunchKernelGGL((panel_gemv_sub_kernel<TS, TC>), dim3((unsigned)((r0 + 63) / 64), (unsigned)nrhs), dim3(256),
                         0, s, LU, ld, (int64_t)0, r0, r0, w, xb, tmp, n);
    else
      hipLaunchKernelGGL(copy_block_kernel<CV>, dim3(1, (unsigned)nrhs), dim3(256), 0, s, xb, tmp, n, r0, w);
  }
  MOM_HIP(hipGetLastError());
  MOM_HIP(hipMemcpy2DAsync(B, ldb * sizeof(CV), tmp, n * sizeof(CV), n * sizeof(CV), nrhs, hipMemcpyDeviceToDevice,
                           s));
  return 0;
}

// b[r] -= sum_{i < w} conj(A[c0 + i, r]) xb[i] for rows [r0, r1): the block
// row c0.. of a triangular factor, conjugate-transposed, read as column
// segments (one wave per output row, four per workgroup)
template <typename TS, typename TC>
__global__ __launch_bounds__(256) void panel_gemv_sub_h_kernel(const typename Cx<TS>::type *__restrict__ A,
                                                                int64_t ld, int64_t r0, int64_t r1, int64_t c0, int w,
                                                                const typename Cx<TC>::type *__restrict__ xb,
                                                                typename Cx<TC>::type *b, int64_t ldb) {
  using CV = typename Cx<TC>::type;
  const int64_t rhs = blockIdx.y;
  // the solved block into its own rows [c0, c0 + w), outside [r0, r1)
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < w; i += 256) b[rhs * ldb + c0 + i] = xb[rhs * NB + i];
  const int lane = threadIdx.x & 63;
  const int64_t r = r0 + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= r1) return;
  const typename Cx<TS>::type *a = A + r * ld + c0;
  const CV *x = xb + rhs * NB;
  CV acc{0, 0};
  for (int i = lane; i < w; i += 64) {
    typename Cx<TS>::type m = a[i];
    m.y = -m.y;
    cmac<TS, TC>(acc, m, x[i]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    acc.x += __shfl_xor(acc.x, o);
    acc.y += __shfl_xor(acc.y, o);
  }
  if (lane == 0) {
    CV *p = b + rhs * ldb + r;
    CV v = *p;
    v.x -= acc.x;
    v.y -= acc.y;
    *p = v;
  }
}

template <typename TV>
__global__ void scatter_rows_kernel(const TV *__restrict__ src, int64_t lds, TV *__restrict__ dst, int64_t ldd,
                                    const int *__restrict__ perm, int64_t n, int64_t nrhs) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * nrhs) return;
  const int64_t i = t % n, r = t / n;
  dst[r * ldd + perm[i]] = src[r * lds + i];
}

// Solve Z^H x = b in place on B with P Z = L U (perm as in getrs):
// Z^H = U^H L^H P, so forward substitution through U^H (diagonal blocks
// Uinv^H), backward through L^H (Linv^H), and the rows scattered back.
template <typename TS, typename TC>
static int getrs_h(hipStream_t s, const typename Cx<TS>::type *LU, int64_t ld, int64_t n, const int *perm,
                   const double2 *Linv, const double2 *Uinv, typename Cx<TC>::type *B, int64_t ldb, int64_t nrhs,
                   typename Cx<TC>::type *tmp, typename Cx<TC>::type *xb) {
  using CV = typename Cx<TC>::type;
  if (nrhs > 65535) return fail(MOM_ERR_ARG, "too many right-hand sides for one adjoint solve");
  MOM_HIP(hipMemcpy2DAsync(tmp, n * sizeof(CV), B, ldb * sizeof(CV), n * sizeof(CV), nrhs, hipMemcpyDeviceToDevice,
                           s));
  const int64_t nblk = (n + NB - 1) / NB;
  for (int64_t b = 0; b < nblk; b++) {
    const int64_t r0 = b * NB;
    const int w = (int)std::min<int64_t>(NB, n - r0);
    hipLaunchKernelGGL((block_apply_kernel<TS, TC, true>), dim3(NB / 8, (unsigned)nrhs), dim3(256), 0, s,
                       Uinv + b * NB * NB, tmp + r0, n, xb, w);
    if (r0 + w < n)
      hipLaunchKernelGGL((panel_gemv_sub_h_kernel<TS, TC>), dim3((unsigned)((n - r0 - w + 3) / 4), (unsigned)nrhs),
                         dim3(256), 0, s, LU, ld, r0 + w, n, r0, w, xb, tmp, n);
    else
      hipLaunchKernelGGL(copy_block_kernel<CV>, dim3(1, (unsigned)nrhs), dim3(256), 0, s, xb, tmp, n, r0, w);
  }
  for (int64_t b = nblk - 1; b >= 0; b--) {
    const int64_t r0 = b * NB;
    const int w = (int)std::min<int64_t>(NB, n - r0);
    hipLaunchKernelGGL((block_apply_kernel<TS, TC, true>), dim3(NB / 8, (unsigned)nrhs), dim3(256), 0, s,
                       Linv + b * NB * NB, tmp + r0, n, xb, w);
    if (r0 > 0)
      hipLaunchKernelGGL((panel_gemv_sub_h_kernel<TS, TC>), dim3((unsigned)((r0 + 3) / 4), (unsigned)nrhs), dim3(256),
                         0, s, LU, ld, (int64_t)0, r0, r0, w, xb, tmp, n);
    else
      hipLaunchKernelGGL(copy_block_kernel<CV>, dim3(1, (unsigned)nrhs), dim3(256), 0, s, xb, tmp, n, r0, w);
  }
  hipLaunchKernelGGL(scatter_rows_kernel<CV>, dim3((unsigned)((n * nrhs + 255) / 256)), dim3(256), 0, s, tmp, n, B,
                     ldb, perm, n, nrhs);
  MOM_HIP(hipGetLastError());
  return 0;
}

}  // namespace mom

// ---------------------------------------------------------------------------
// Complex128 matrix-vector kernels for the residual and the scaling.
namespace mom {

constexpr int kChunks = 32;  // column chunks of the chunked row kernels

// part[chunk][r] = sum_{c in chunk} Z[r, c] x[c]
__global__ __launch_bounds__(256) void zgemv_part_kernel(const double2 *__restrict__ Z, int64_t ld, int64_t n,
                                                          const double2 *__restrict__ x, double2 *part) {
  __shared__ double2 xs[2048];
  const int64_t cw = (n + kChunks - 1) / kChunks;
  const int64_t c0 = blockIdx.y * cw, c1 = std::min<int64_t>(n, c0 + cw);
  for (int64_t c = c0 + threadIdx.x; c < c1; c += 256) xs[c - c0] = x[c];
  __syncthreads();
  const int64_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  double ar = 0, ai = 0;
  const double2 *z = Z + c0 * ld + r;
  for (int64_t c = c0; c < c1; c++, z += ld) {
    const double2 m = *z, v = xs[c - c0];
    ar += m.x * v.x - m.y * v.y;
    ai += m.x * v.y + m.y * v.x;
  }
  part[blockIdx.y * n + r] = make_double2(ar, ai);
}

// y = b - sum_chunks part (b may be NULL: y = -Z x)
__global__ void zgemv_reduce_kernel(const double2 *__restrict__ part, int64_t n, const double2 *__restrict__ b,
                                    double2 *y) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  double ar = 0, ai = 0;
  for (int k = 0; k < kChunks; k++) {
    ar += part[k * n + r].x;
    ai += part[k * n + r].y;
  }
  double2 o = b ? b[r] : make_double2(0, 0);
  y[r] = make_double2(o.x - ar, o.y - ai);
}

// y[r] = b[r] - sum_c conj(Z[c, r]) x[c] (b may be NULL): Z^H x, one wave
// per column of Z (contiguous reads), lane partial sums in a fixed order
__global__ __launch_bounds__(256) void zgemv_h_kernel(const double2 *__restrict__ Z, int64_t ld, int64_t n,
                                                       const double2 *__restrict__ x, const double2 *__restrict__ b,
                                                       double2 *y) {
  const int lane = threadIdx.x & 63;
  const int64_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n) return;
  const double2 *z = Z + r * ld;
  double ar = 0, ai = 0;
  for (int64_t c = lane; c < n; c += 64) {
    const double2 m = z[c], v = x[c];
    ar += m.x * v.x + m.y * v.y;
    ai += m.x * v.y - m.y * v.x;
  }
  for (int o = 32; o > 0; o >>= 1) {
    ar += __shfl_xor(ar, o);
    ai += __shfl_xor(ai, o);
  }
  if (lane == 0) {
    const double2 o = b ? b[r] : make_double2(0, 0);
    y[r] = make_double2(o.x - ar, o.y - ai);
  }
}

// out[r] = sum_c |Z[c, r]| (column sums: ||Z||_1 = ||Z^H||_inf)
__global__ __launch_bounds__(256) void col_abssum_kernel(const double2 *__restrict__ Z, int64_t ld, int64_t n,
                                                          double *out) {
  const int lane = threadIdx.x & 63;
  const int64_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n) return;
  const double2 *z = Z + r * ld;
  double a = 0;
  for (int64_t c = lane; c < n; c += 64) a += std::sqrt(z[c].x * z[c].x + z[c].y * z[c].y);
  for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
  if (lane == 0) out[r] = a;
}

// part[chunk][r] = (max_c |Z[r,c]|, sum_c |Z[r,c]|) over the chunk
__global__ __launch_bounds__(256) void row_stats_part_kernel(const double2 *__restrict__ Z, int64_t ld, int64_t n,
                                                              double2 *part) {
  const int64_t cw = (n + kChunks - 1) / kChunks;
  const int64_t c0 = blockIdx.y * cw, c1 = std::min<int64_t>(n, c0 + cw);
  const int64_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  double mx = 0, sm = 0;
  const double2 *z = Z + c0 * ld + r;
  for (int64_t c = c0; c < c1; c++, z += ld) {
    const double2 m = *z;
    const double a = sqrt(m.x * m.x + m.y * m.y);
    mx = fmax(mx, a);
    sm += a;
  }
  part[blockIdx.y * n + r] = make_double2(mx, sm);
}

// rscale[r] = 2^-ilogb(rowmax); rowsum[r] = sum_c |Z[r,c]|
__global__ void row_stats_reduce_kernel(const double2 *__restrict__ part, int64_t n, double *rscale,
                                        double *rowsum) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  double mx = 0, sm = 0;
  for (int k = 0; k < kChunks; k++) {
    mx = fmax(mx, part[k * n + r].x);
    sm += part[k * n + r].y;
  }
  rscale[r] = (mx > 0 && isfinite(mx)) ? ldexp(1.0, -ilogb(mx)) : 1.0;
  rowsum[r] = sm;
}

// cscale[c] = 2^-ilogb(max_r rscale[r] |Z[r,c]|), one workgroup per column
__global__ __launch_bounds__(256) void col_scale_kernel(const double2 *__restrict__ Z, int64_t ld, int64_t n,
                                                         const double *__restrict__ rscale, double *cscale) {
  __shared__ double sm[256];
  const int64_t c = blockIdx.x;
  double mx = 0;
  const double2 *z = Z + c * ld;
  for (int64_t r = threadIdx.x; r < n; r += 256) {
    const double2 m = z[r];
    mx = fmax(mx, rscale[r] * sqrt(m.x * m.x + m.y * m.y));
  }
  sm[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) cscale[c] = (sm[0] > 0 && isfinite(sm[0])) ? ldexp(1.0, -ilogb(sm[0])) : 1.0;
}

// Z32[r,c] = fl32(rscale[r] Z[r,c] cscale[c])
__global__ __launch_bounds__(256) void to_fp32_kernel(const double2 *__restrict__ Z, int64_t ld, int64_t n,
                                                       const double *__restrict__ rscale,
                                                       const double *__restrict__ cscale, float2 *Z32) {
  const int64_t c = blockIdx.y;
  const double cs = cscale[c];
  const double2 *z = Z + c * ld;
  float2 *o = Z32 + c * n;
  for (int64_t r = blockIdx.x * 256 + threadIdx.x; r < n; r += (int64_t)gridDim.x * 256) {
    const double s = rscale[r] * cs;
    const double2 m = z[r];
    o[r] = make_float2((float)(m.x * s), (float)(m.y * s));
  }
}

// y = d .* v (diagonal scaling, complex128)
__global__ void dscale_kernel(const double2 *__restrict__ v, const double *__restrict__ d, int64_t n, double2 *y) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  y[r] = make_double2(v[r].x * d[r], v[r].y * d[r]);
}

// ---------------------------------------------------------------------------
// Many right-hand sides: triangular solves and refinement as matrix
// products (rocBLAS ZGEMM) on n x nrhs blocks, instead of one vector at a
// time.

// y[:, j] = d .* v[:, j]
__global__ void dscale_cols_kernel(const double2 *__restrict__ v, int64_t ldv, const double *__restrict__ d,
                                   int64_t n, double2 *y, int64_t ldy) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, j = blockIdx.y;
  if (r >= n) return;
  const double2 a = v[j * ldv + r];
  y[j * ldy + r] = make_double2(a.x * d[r], a.y * d[r]);
}

// x[:, j] += d .* z[:, j]
__global__ void scaled_add_cols_kernel(double2 *x, int64_t ldx, const double *__restrict__ d,
                                       const double2 *__restrict__ z, int64_t ldz, int64_t n) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, j = blockIdx.y;
  if (r >= n) return;
  const double2 a = z[j * ldz + r];
  double2 &o = x[j * ldx + r];
  o.x += a.x * d[r];
  o.y += a.y * d[r];
}

// out[j] = max_r |v[r, j]| (one workgroup per column)
__global__ __launch_bounds__(256) void col_absmax_kernel(const double2 *__restrict__ v, int64_t ld, int64_t n,
                                                          double *out) {
  __shared__ double sm[4];
  const int64_t j = blockIdx.x;
  double m = 0;
  for (int64_t r = threadIdx.x; r < n; r += 256) {
    const double2 a = v[j * ld + r];
    m = fmax(m, sqrt(a.x * a.x + a.y * a.y));
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) out[j] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
}

// columns j with keep[j] == 0 are set to zero
__global__ void mask_cols_kernel(double2 *v, int64_t ld, int64_t n, const int *__restrict__ keep) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, j = blockIdx.y;
  if (r >= n || keep[j]) return;
  v[j * ld + r] = make_double2(0.0, 0.0);
}

// P[r, c] = A[r, c] widened to complex128, rows x w
__global__ void widen_panel_kernel(const float2 *__restrict__ A, int64_t ld, int64_t rows, double2 *P,
                                   int64_t ldp) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, c = blockIdx.y;
  if (r >= rows) return;
  const float2 a = A[c * ld + r];
  P[c * ldp + r] = make_double2(a.x, a.y);
}

// Solve L U X = P B in place on B (n x nrhs, ldb), complex128 arithmetic
// with the factors stored in TS: per 256-row block the diagonal-block
// inverse (complex128) times the block, then the panel below (forward) or
// above (backward) times the solved block subtracted from the rest; a
// complex64 panel is first widened into P (n x NB).  T is n x nrhs and W
// NB x nrhs scratch.
template <typename TS>
static int getrs_blocked(mom_ctx *ctx, const typename Cx<TS>::type *LU, int64_t ld, int64_t n, const int *perm,
                         const double2 *Linv, const double2 *Uinv, double2 *B, int64_t ldb, int64_t nrhs,
                         double2 *T, double2 *W, double2 *P) {
  hipStream_t s = ctx->stream;
  const Lane L = main_lane(ctx);
  hipLaunchKernelGGL(gather_rows_kernel<double2>, dim3((unsigned)((n * nrhs + 255) / 256)), dim3(256), 0, s, B, ldb,
                     T, n, perm, n, nrhs);
  MOM_HIP(hipGetLastError());
  auto panel = [&](int64_t r0, int64_t rows, int64_t c0, int64_t w, const double2 **A, int64_t *lda) -> int {
    if constexpr (std::is_same<TS, double>::value) {
      *A = LU + c0 * ld + r0;
      *lda = ld;
    } else {
      hipLaunchKernelGGL(widen_panel_kernel, dim3((unsigned)((rows + 255) / 256), (unsigned)w), dim3(256), 0, s,
                         LU + c0 * ld + r0, ld, rows, P, rows);
      MOM_HIP(hipGetLastError());
      *A = P;
      *lda = rows;
    }
    return 0;
  };
  const int64_t nblk = (n + NB - 1) / NB;
  for (int64_t b = 0; b < nblk; b++) {
    const int64_t r0 = b * NB, w = std::min<int64_t>(NB, n - r0);
    MOM_TRY(zgemm(L, w, nrhs, w, 1.0, true, Linv + b * NB * NB, NB, T + r0, n, W, NB));
    MOM_HIP(hipMemcpy2DAsync(T + r0, n * sizeof(double2), W, NB * sizeof(double2), w * sizeof(double2), nrhs,
                             hipMemcpyDeviceToDevice, s));
    if (r0 + w < n) {
      const double2 *A;
      int64_t lda;
      MOM_TRY(panel(r0 + w, n - r0 - w, r0, w, &A, &lda));
      MOM_TRY(zgemm(L, n - r0 - w, nrhs, w, -1.0, false, A, lda, W, NB, T + r0 + w, n));
    }
  }
  for (int64_t b = nblk - 1; b >= 0; b--) {
    const int64_t r0 = b * NB, w = std::min<int64_t>(NB, n - r0);
    MOM_TRY(zgemm(L, w, nrhs, w, 1.0, true, Uinv + b * NB * NB, NB, T + r0, n, W, NB));
    MOM_HIP(hipMemcpy2DAsync(T + r0, n * sizeof(double2), W, NB * sizeof(double2), w * sizeof(double2), nrhs,
                             hipMemcpyDeviceToDevice, s));
    if (r0 > 0) {
      const double2 *A;
      int64_t lda;
      MOM_TRY(panel(0, r0, r0, w, &A, &lda));
      MOM_TRY(zgemm(L, r0, nrhs, w, -1.0, false, A, lda, W, NB, T, n));
    }
  }
  MOM_HIP(hipMemcpy2DAsync(B, ldb * sizeof(double2), T, n * sizeof(double2), n * sizeof(double2), nrhs,
                           hipMemcpyDeviceToDevice, s));
  return 0;
}

// T[r, j] = (P R v)[r, j] rounded to complex64: v scaled by the row
// equilibration d and permuted (row r of the result is row perm[r] of d v)
__global__ void scale_gather_round_kernel(const double2 *__restrict__ v, int64_t ldv, const double *__restrict__ d,
                                          const int *__restrict__ perm, int64_t n, float2 *__restrict__ T) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, j = blockIdx.y;
  if (r >= n) return;
  const int64_t pr = perm[r];
  const double2 a = v[j * ldv + pr];
  T[j * n + r] = make_float2((float)(a.x * d[pr]), (float)(a.y * d[pr]));
}

// x[:, j] (+)= d .* T[:, j] (T complex64 widened)
__global__ void widen_scale_store_kernel(double2 *x, int64_t ldx, const double *__restrict__ d,
                                         const float2 *__restrict__ T, int64_t n, int add) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, j = blockIdx.y;
  if (r >= n) return;
  const float2 t = T[j * n + r];
  double2 &o = x[j * ldx + r];
  const double2 v = make_double2((double)t.x * d[r], (double)t.y * d[r]);
  if (add) {
    o.x += v.x;
    o.y += v.y;
  } else {
    o = v;
  }
}

// rows [r0, r0 + w) of T (complex64, n x nrhs) widened into W (NB x nrhs),
// and back rounded
__global__ void widen_rows_kernel(const float2 *__restrict__ T, int64_t n, int64_t r0, int w, double2 *W) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (r >= w) return;
  const float2 t = T[j * n + r0 + r];
  W[(int64_t)j * NB + r] = make_double2(t.x, t.y);
}
__global__ void round_rows_kernel(const double2 *__restrict__ W, int64_t n, int64_t r0, int w, float2 *T) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (r >= w) return;
  const double2 v = W[(int64_t)j * NB + r];
  T[j * n + r0 + r] = make_float2((float)v.x, (float)v.y);
}

// L U T = T for a block of right-hand sides in complex64 (T, n x nrhs,
// leading dimension n, already permuted): per 256-row block the complex128
// diagonal-block inverse times the block (in FP64, then rounded), then the
// complex64 panel below (forward) or above (backward) times the solved
// block subtracted from the rest on the complex64 matrix cores (the
// trailing-update kernels).  Used by the block refinement, whose steps only
// need a preconditioner as accurate as the complex64 factors themselves;
// the single-column GMRES keeps the exact FP64 application.  Wa, Wb: NB x
// nrhs complex128 scratch.
static int getrs_blocked32(mom_ctx *ctx, const float2 *LU, int64_t ld, int64_t n, const double2 *Linv,
                           const double2 *Uinv, float2 *T, int64_t nrhs, double2 *Wa, double2 *Wb) {
  hipStream_t s = ctx->stream;
  const Lane L = main_lane(ctx);
  const int64_t nblk = (n + NB - 1) / NB;
  auto diag = [&](const double2 *Inv, int64_t r0, int w) -> int {
    const dim3 g((unsigned)((w + 255) / 256), (unsigned)nrhs);
    hipLaunchKernelGGL(widen_rows_kernel, g, dim3(256), 0, s, T, n, r0, w, Wa);
    MOM_TRY(gemm_generic<double>(s, Inv, NB, Wa, NB, Wb, NB, w, nrhs, w, 1.0, true));
    hipLaunchKernelGGL(round_rows_kernel, g, dim3(256), 0, s, Wb, n, r0, w, T);
    MOM_HIP(hipGetLastError());
    return 0;
  };
  for (int64_t b = 0; b < nblk; b++) {
    const int64_t r0 = b * NB, w = std::min<int64_t>(NB, n - r0);
    MOM_TRY(diag(Linv + b * NB * NB, r0, (int)w));
    if (r0 + w < n)
      MOM_TRY(gemm_sub<float>(L, LU + r0 * ld + r0 + w, ld, T + r0, n, T + r0 + w, n, n - r0 - w, nrhs, w));
  }
  for (int64_t b = nblk - 1; b >= 0; b--) {
    const int64_t r0 = b * NB, w = std::min<int64_t>(NB, n - r0);
    MOM_TRY(diag(Uinv + b * NB * NB, r0, (int)w));
    if (r0 > 0) MOM_TRY(gemm_sub<float>(L, LU + r0 * ld, ld, T + r0, n, T, n, r0, nrhs, w));
  }
  return 0;
}

template <typename T>
static int diag_inverses(hipStream_t s, const typename Cx<T>::type *LU, int64_t ld, int64_t n, double2 *Linv,
                         double2 *Uinv) {
  const int64_t nblk = (n + NB - 1) / NB;
  hipLaunchKernelGGL(diag_inverse_kernel<T>, dim3((unsigned)nblk, NB / kInvCols, 2), dim3(256), 0, s, LU, ld, n, Linv,
                     Uinv);
  MOM_HIP(hipGetLastError());
  return 0;
}

static double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace mom

namespace mom {
// The complex64 factorisation with tournament pivoting (getrf and the
// diagonal-block inverses) as a captured HIP graph: its launches depend only
// on n and on the buffers they touch, so a system captures it on its first
// factorisation and replays it on later ones.  `key` holds everything the
// captured launches baked in (buffer addresses, the look-ahead stream, n and
// the process-wide switches); any change captures again.
struct FactorGraph {
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  std::vector<uintptr_t> key;
  void reset() {
    if (x) (void)hipGraphExecDestroy(x);
    if (g) (void)hipGraphDestroy(g);
    x = nullptr;
    g = nullptr;
    key.clear();
  }
  ~FactorGraph() { reset(); }
};
}  // namespace mom

struct mom_system {
  mom_ctx *ctx = nullptr;
  int64_t N = 0;  // nedges_total
  int64_t n = 0;  // 2N
  int precision = MOM_PREC_FP64;       // precision of the current factors
  int precision_req = MOM_PREC_FP64;   // asked for at creation or by mom_system_set_precision (set by every fill)
  bool fell_back = false;              // the last block solve moved a mixed system to FP64 factors
  bool filled = false, factored = false;
  mom::DevBuf Z64, Z32, ipiv, Linv, Uinv, rscale, cscale, rowsum, part, pivwork;
  std::vector<int> perm;
  double znorm_inf = 0;
  double znorm_one = 0;  // ||Z||_1 of the unfactored complex128 matrix (adjoint refinement; 0: not yet)
  double t_fill = 0, t_factor = 0, t_solve = 0;
  mom::PhaseClock clock;
  int info = 0;
  // cooperative panel factorisation: synchronisation words (zeroed when
  // allocated), next round tag, co-resident grid limit per precision
  mom::DevBuf panel_sync;
  unsigned panel_tag = 1;
  int n_minv = 0, n_matvec = 0;  // preconditioner applications / FP64 products of the last refinement
  double be_hist[8] = {-1, -1, -1, -1, -1, -1, -1, -1};  // backward error after each outer step
  int coop_gmax[2] = {-1, -1};
  bool coop_refused = false;  // a cooperative launch was refused: panels by one launch per column
  bool coop_off = false;      // mom_system_set_cooperative(0): panels by one launch per column
  int pivoting = 0;  // complex64 panels: 0 partial pivoting, 1 tournament (mom_system_set_pivoting)
  mom::FactorGraph fgraph;  // the tournament factorisation, captured (see factor_graph)
};

namespace mom {

static int ensure_pivwork(mom_system *S, PivotWork *w) {
  const int64_t mb = (S->n + 255) / 256 + 1;
  const size_t rowbytes = 2 * mb * IB * sizeof(double2);
  const size_t planb = ((1 + 4 * NB) * sizeof(int) + 255) / 256 * 256;
  const size_t tselb = (2 * mb * IB * sizeof(int) + 255) / 256 * 256;
  const size_t bytes = rowbytes + 2 * mb * (sizeof(double) + sizeof(int)) + 64 + planb + tselb +
                       IB * IB * sizeof(float2);
  if (S->pivwork.bytes < bytes) {
    int rc = S->pivwork.alloc(bytes);
    if (rc) return rc;
  }
  char *base = S->pivwork.as<char>();
  w->row = base;
  w->val = (double *)(base + rowbytes);
  w->idx = (int *)(base + rowbytes + 2 * mb * sizeof(double));
  w->info = (int *)(base + rowbytes + 2 * mb * (sizeof(double) + sizeof(int)));
  w->plan = (int *)(base + rowbytes + 2 * mb * (sizeof(double) + sizeof(int)) + 64);
  w->tsel = (int *)(base + rowbytes + 2 * mb * (sizeof(double) + sizeof(int)) + 64 + planb);
  w->lu_top = (float2 *)(base + rowbytes + 2 * mb * (sizeof(double) + sizeof(int)) + 64 + planb + tselb);
  w->mb = (int)mb;
  w->tournament = (S->pivoting == 1 || g_tournament) ? 1 : 0;
  MOM_HIP(hipMemsetAsync(w->info, 0, sizeof(int), S->ctx->stream));
  return 0;
}

static int g_panel_mode = 0;
constexpr int kBeHistory = 8;  // 0: cooperative inner blocks when they fit, 1: one launch per column

// Sets up the cooperative panel path for precision T: the synchronisation
// buffer (allocated once per system and zeroed) and the largest grid whose
// workgroups are all resident at once.
template <typename T>
static int panel_coop(mom_system *S, PanelCoop *c) {
  const int pi = std::is_same<T, double>::value ? 1 : 0;
  if (S->coop_gmax[pi] < 0) {
    int dev = 0, coop = 0, nb = 0;
    MOM_HIP(hipGetDevice(&dev));
    MOM_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    hipDeviceProp_t prop;
    MOM_HIP(hipGetDeviceProperties(&prop, dev));
    MOM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)panel_block_kernel<T>, 256, 0));
    S->coop_gmax[pi] = coop ? std::min(kPanelMaxGroups, nb * prop.multiProcessorCount) : 0;
  }
  c->gmax = (g_panel_mode == 1 || S->coop_off) ? 0 : S->coop_gmax[pi];
  c->refused = &S->coop_refused;
  if (c->gmax == 0) return 0;
  const size_t cand = sizeof(u64) * 2 * kPanelMaxGroups * kCandWords, diag = sizeof(u64) * 2 * kRowWords;
  if (!S->panel_sync.ptr) {
    MOM_TRY(S->panel_sync.alloc(cand + diag + 64));
    MOM_HIP(hipMemsetAsync(S->panel_sync.ptr, 0, cand + diag + 64, S->ctx->stream));
    S->panel_tag = 1;
  }
  char *base = S->panel_sync.as<char>();
  c->sync.cand = (u64 *)base;
  c->sync.diag = (u64 *)(base + cand);
  c->sync.err = (int *)(base + cand + diag);
  c->sync.gmax = kPanelMaxGroups;
  if (S->panel_tag > 0xF0000000u) {  // tags are about to wrap: start over from clean words
    MOM_HIP(hipMemsetAsync(S->panel_sync.ptr, 0, cand + diag, S->ctx->stream));
    S->panel_tag = 1;
  }
  c->tag = &S->panel_tag;
  // a timeout of an earlier factorisation must not end this one's polls
  MOM_HIP(hipMemsetAsync(c->sync.err, 0, sizeof(int), S->ctx->stream));
  return 0;
}

// Whether factorisations with tournament pivoting replay a captured graph
// (process-wide; the test hook mom_test_factor_graph switches it), and how
// many captures and replays there were.
static int g_factor_graph = 1;
static std::atomic<int64_t> g_graph_counts[2] = {0, 0};

// Factorisations whose panels are cooperative launches hold this lock while
// their launches are in flight: two such grids made by different host
// threads (systems of different contexts) could each fill the device with
// workgroups that wait for the rest of their own grid.  Launches of other
// kinds from other threads may run beside one cooperative grid, whose
// missing workgroups then start as soon as those finish.
static std::mutex g_coop_mu;

// getrf and the diagonal-block inverses of the complex64 factors with
// tournament pivoting, from the system's captured graph.  Every buffer the
// launches touch is sized beforehand (no allocation while capturing), and
// the capture is thread-local, so other host threads may keep allocating.
static int factor_graph(mom_system *S, float2 *A, const PivotWork &w, const PanelCoop &coop) {
  mom_ctx *ctx = S->ctx;
  hipStream_t s = ctx->stream;
  const int64_t n = S->n;
  const bool la = g_lookahead != 0 && n > NB;
  Lane X;
  if (la) MOM_TRY(aux_lane(ctx, &X));
  const size_t bt = sizeof(float2) * NB * (size_t)((n + kDmaBN - 1) / kDmaBN * kDmaBN);
  if (ctx->panel_t.bytes < bt) MOM_TRY(ctx->panel_t.alloc(bt));
  if (la && ctx->panel_t2.bytes < bt) MOM_TRY(ctx->panel_t2.alloc(bt));
  if (n > NB && ctx->ublk.bytes < sizeof(float2) * NB * NB) MOM_TRY(ctx->ublk.alloc(sizeof(float2) * NB * NB));
  const std::vector<uintptr_t> key = {
      (uintptr_t)A, (uintptr_t)S->ipiv.ptr, (uintptr_t)S->pivwork.ptr, (uintptr_t)S->Linv.ptr,
      (uintptr_t)S->Uinv.ptr, (uintptr_t)ctx->panel_t.ptr, (uintptr_t)ctx->panel_t2.ptr, (uintptr_t)ctx->ublk.ptr,
      (uintptr_t)ctx->aux, (uintptr_t)ctx->fork_ev, (uintptr_t)ctx->join_ev, (uintptr_t)n, (uintptr_t)g_lookahead,
      (uintptr_t)g_u_subst};
  FactorGraph &G = S->fgraph;
  if (!G.x || G.key != key) {
    G.reset();
    MOM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = getrf<float>(ctx, A, n, n, S->ipiv.as<int>(), w, nullptr, coop);
    if (!rc) rc = diag_inverses<float>(s, A, n, n, S->Linv.as<double2>(), S->Uinv.as<double2>());
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      return rc ? rc : fail(MOM_ERR_HIP, std::string("hipStreamEndCapture(factor): ") + hipGetErrorString(e));
    }
    G.g = g;
    MOM_HIP(hipGraphInstantiate(&G.x, G.g, nullptr, nullptr, 0));
    G.key = key;
    g_graph_counts[0]++;
  } else {
    g_graph_counts[1]++;
  }
  MOM_HIP(hipGraphLaunch(G.x, s));
  return 0;
}

// Factor the n x n matrix held by the system (complex128 in place, or the
// scaled complex64 copy), build the permutation and diagonal-block inverses.
template <typename T>
static int factor_impl(mom_system *S, typename Cx<T>::type *A) {
  hipStream_t s = S->ctx->stream;
  const int64_t n = S->n;
  PivotWork w;
  int rc;
  if ((rc = ensure_pivwork(S, &w))) return rc;
  if (S->clock.enabled) S->clock.reset();
  PanelCoop coop;
  if ((rc = panel_coop<T>(S, &coop))) return rc;
  const int64_t nblk = (n + NB - 1) / NB;
  const size_t inv_bytes = sizeof(double2) * nblk * NB * NB;  // complex128 in both precisions
  if (S->Linv.bytes < inv_bytes && ((rc = S->Linv.alloc(inv_bytes)) || (rc = S->Uinv.alloc(inv_bytes)))) return rc;
  bool graphed = false;
  if constexpr (std::is_same<T, float>::value) {
    // tournament panels make no cooperative launches and no host reads
    // inside getrf; event timing and the E = I - inv(L11) variant keep the
    // launches eager
    if (w.tournament && g_factor_graph && !S->clock.enabled && g_u_subst != 0) {
      if ((rc = factor_graph(S, A, w, coop))) return rc;
      graphed = true;
    }
  }
  std::unique_lock<std::mutex> coop_lock(g_coop_mu, std::defer_lock);
  if (!graphed) {
    if (coop.gmax > 0 && !w.tournament) coop_lock.lock();  // released once the pivots have been read back
    if ((rc = getrf<T>(S->ctx, A, n, n, S->ipiv.as<int>(), w, &S->clock, coop))) {
      if (coop_lock.owns_lock()) (void)hipStreamSynchronize(s);
      return rc;
    }
    if ((rc = diag_inverses<T>(s, A, n, n, S->Linv.as<double2>(), S->Uinv.as<double2>()))) {
      if (coop_lock.owns_lock()) (void)hipStreamSynchronize(s);
      return rc;
    }
  }
  std::vector<int> ip(n);
  int info = 0;
  MOM_TRY(copy_d2h(S->ctx, ip.data(), S->ipiv.ptr, sizeof(int) * n));
  MOM_TRY(copy_d2h(S->ctx, &info, w.info, sizeof(int)));
  if (coop.gmax > 0) {
    int timed_out = 0;
    MOM_TRY(copy_d2h(S->ctx, &timed_out, coop.sync.err, sizeof(int)));
    if (timed_out) return fail(MOM_ERR_HIP, "panel factorisation: workgroup exchange timed out");
  }
  if (S->clock.enabled && (rc = S->clock.collect())) return rc;
  S->perm.resize(n);
  for (int64_t i = 0; i < n; i++) S->perm[i] = (int)i;
  for (int64_t k = 0; k < n; k++) {
    if (ip[k] - 1 < k || ip[k] - 1 >= n) return fail(MOM_ERR_HIP, "factorisation produced an invalid pivot row");
    std::swap(S->perm[k], S->perm[ip[k] - 1]);
  }
  S->info = info;
  return 0;
}

static int upload_perm(mom_system *S, DevBuf &dperm) {
  int rc = dperm.alloc(sizeof(int) * S->n);
  if (rc) return rc;
  MOM_TRY(copy_h2d(S->ctx, dperm.ptr, S->perm.data(), sizeof(int) * S->n));
  return 0;
}

// Mixed precision solve of one right-hand side b (host, complex128) into x.
struct IrWork {
  DevBuf dperm, v64, w64, t64, tmp64, xb64, b64, x64;
  bool adj = false;  // refine Z^H x = b (preconditioner R M^-H C, residual b - Z^H x)
};

static int apply_minv(mom_system *S, IrWork &W, const std::vector<cd> &v, std::vector<cd> &y) {
  // M^-1 v = C U^-1 L^-1 P R v with the complex64 factors of R Z C, applied
  // in complex128 arithmetic
  hipStream_t s = S->ctx->stream;
  S->n_minv++;
  const int64_t n = S->n;
  MOM_TRY(copy_h2d(S->ctx, W.v64.ptr, v.data(), sizeof(double2) * n));
  const unsigned g = (unsigned)((n + 255) / 256);
  // adjoint: Z^-H = R M^-H C with M = R Z C
  hipLaunchKernelGGL(dscale_kernel, dim3(g), dim3(256), 0, s, W.v64.as<double2>(),
                     (W.adj ? S->cscale : S->rscale).as<double>(), n, W.t64.as<double2>());
  int rc = W.adj ? getrs_h<float, double>(s, S->Z32.as<float2>(), n, n, W.dperm.as<int>(), S->Linv.as<double2>(),
                                         S->Uinv.as<double2>(), W.t64.as<double2>(), n, 1, W.tmp64.as<double2>(),
                                         W.xb64.as<double2>())
                 : getrs<float, double>(s, S->Z32.as<float2>(), n, n, W.dperm.as<int>(), S->Linv.as<double2>(),
                                        S->Uinv.as<double2>(), W.t64.as<double2>(), n, 1, W.tmp64.as<double2>(),
                                        W.xb64.as<double2>());
  if (rc) return rc;
  hipLaunchKernelGGL(dscale_kernel, dim3(g), dim3(256), 0, s, W.t64.as<double2>(),
                     (W.adj ? S->rscale : S->cscale).as<double>(), n, W.w64.as<double2>());
  MOM_HIP(hipGetLastError());
  y.resize(n);
  return copy_d2h(S->ctx, y.data(), W.w64.ptr, sizeof(double2) * n);
}

// y = b - Z x (b may be NULL: y = -Z x); x, y on the host
static int residual(mom_system *S, IrWork &W, const std::vector<cd> *b, const std::vector<cd> &x,
                    std::vector<cd> &y) {
  hipStream_t s = S->ctx->stream;
  S->n_matvec++;
  const int64_t n = S->n;
  MOM_TRY(copy_h2d(S->ctx, W.x64.ptr, x.data(), sizeof(double2) * n));
  if (b) MOM_TRY(copy_h2d(S->ctx, W.b64.ptr, b->data(), sizeof(double2) * n));
  if (W.adj) {
    hipLaunchKernelGGL(zgemv_h_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, S->Z64.as<double2>(), n, n,
                       W.x64.as<double2>(), b ? W.b64.as<double2>() : nullptr, W.w64.as<double2>());
  } else {
    hipLaunchKernelGGL(zgemv_part_kernel, dim3((unsigned)((n + 255) / 256), kChunks), dim3(256), 0, s,
                       S->Z64.as<double2>(), n, n, W.x64.as<double2>(), S->part.as<double2>());
    hipLaunchKernelGGL(zgemv_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       S->part.as<double2>(), n, b ? W.b64.as<double2>() : nullptr, W.w64.as<double2>());
  }
  MOM_HIP(hipGetLastError());
  y.resize(n);
  return copy_d2h(S->ctx, y.data(), W.w64.ptr, sizeof(double2) * n);
}

static double inf_norm(const std::vector<cd> &v) {
  double m = 0;
  for (const cd &c : v) m = std::max(m, std::abs(c));
  return m;
}

static cd dotc(const std::vector<cd> &a, const std::vector<cd> &b) {
  cd s = 0;
  for (size_t i = 0; i < a.size(); i++) s += std::conj(a[i]) * b[i];
  return s;
}

// Iterative refinement for one right-hand side, x0 = M^-1 b with the FP32
// factors as M.  Outer loop: r = b - Z x in FP64.  While each step lowers
// the backward error at least tenfold the correction is plain refinement,
// x += M^-1 r; otherwise restarted GMRES on M^-1 Z d = M^-1 r computes it.
static int gmres_ir(mom_system *S, IrWork &W, const std::vector<cd> &b, std::vector<cd> &x, double tol, int maxit,
                    int *iters, double *berr) {
  const int64_t n = S->n;
  int rc;
  S->n_minv = S->n_matvec = 0;
  for (int i = 0; i < kBeHistory; i++) S->be_hist[i] = -1;
  if ((rc = apply_minv(S, W, b, x))) return rc;
  const double bn = inf_norm(b);
  const int m = 20;
  std::vector<std::vector<cd>> V(m + 1);
  std::vector<cd> r, z, w, H((m + 1) * m), g(m + 1), cs(m), sn(m);
  double be = 0, prev = 0;
  bool plain = true;
  int it = 0;
  for (;; it++) {
    if ((rc = residual(S, W, &b, x, r))) return rc;
    be = inf_norm(r) / ((W.adj ? S->znorm_one : S->znorm_inf) * inf_norm(x) + bn);
    if (it < kBeHistory) S->be_hist[it] = be;
    if (be <= tol || it >= maxit) break;
    if (plain && prev > 0 && be > 0.1 * prev) plain = false;
    prev = be;
    if (plain) {
      if ((rc = apply_minv(S, W, r, z))) return rc;
      for (int64_t j = 0; j < n; j++) x[j] += z[j];
      continue;
    }
    // reduce the correction's residual enough for the next outer step to
    // converge: the preconditioned residual GMRES monitors only estimates
    // the backward error, so aim two decades below the target
    const double inner_tol = std::min(1e-2, std::max(1e-14, 0.01 * tol / be));
    if ((rc = apply_minv(S, W, r, z))) return rc;
    double beta = std::sqrt(std::real(dotc(z, z)));
    if (beta == 0) break;
    V[0].resize(n);
    for (int64_t i = 0; i < n; i++) V[0][i] = z[i] / beta;
    std::fill(g.begin(), g.end(), cd(0));
    g[0] = beta;
    int k = 0;
    for (; k < m; k++) {
      // w = M^-1 Z v_k  (residual() returns -Z v)
      if ((rc = residual(S, W, nullptr, V[k], w))) return rc;
      for (auto &c : w) c = -c;
      if ((rc = apply_minv(S, W, w, z))) return rc;
      for (int i = 0; i <= k; i++) {
        cd h = dotc(V[i], z);
        H[i * m + k] = h;
        for (int64_t j = 0; j < n; j++) z[j] -= h * V[i][j];
      }
      double hn = std::sqrt(std::real(dotc(z, z)));
      H[(k + 1) * m + k] = hn;
      V[k + 1].resize(n);
      for (int64_t j = 0; j < n; j++) V[k + 1][j] = hn > 0 ? z[j] / hn : cd(0);
      for (int i = 0; i < k; i++) {
        cd a = H[i * m + k], c2 = H[(i + 1) * m + k];
        H[i * m + k] = std::conj(cs[i]) * a + std::conj(sn[i]) * c2;
        H[(i + 1) * m + k] = -sn[i] * a + cs[i] * c2;
      }
      cd a = H[k * m + k], c2 = H[(k + 1) * m + k];
      double den = std::sqrt(std::norm(a) + std::norm(c2));
      cs[k] = den > 0 ? a / den : cd(1);
      sn[k] = den > 0 ? c2 / den : cd(0);
      H[k * m + k] = std::conj(cs[k]) * a + std::conj(sn[k]) * c2;
      H[(k + 1) * m + k] = 0;
      cd gk = g[k];
      g[k] = std::conj(cs[k]) * gk;
      g[k + 1] = -sn[k] * gk;
      if (std::abs(g[k + 1]) <= inner_tol * beta || hn == 0) {
        k++;
        break;
      }
    }
    // back substitution H y = g, x += V y
    std::vector<cd> y(k);
    for (int i = k - 1; i >= 0; i--) {
      cd acc = g[i];
      for (int j = i + 1; j < k; j++) acc -= H[i * m + j] * y[j];
      y[i] = acc / H[i * m + i];
    }
    for (int i = 0; i < k; i++)
      for (int64_t j = 0; j < n; j++) x[j] += y[i] * V[i][j];
  }
  if (iters) *iters = it;
  if (berr) *berr = be;
  return be <= tol ? 0 : fail(MOM_ERR_NOCONV, "GMRES-IR did not reach the backward-error tolerance");
}

// From this many right-hand sides on, solves run on the whole block at once.
constexpr int64_t kBatchRhs = 4;
// Block refinement: 1 takes the FP64-factor branch whenever a column is
// above the tolerance after the first residual (test hook), 0 decides by
// the cost estimate.
static int g_force_fp64_fallback = 0;

static int ir_alloc(mom_system *S, IrWork &W);

// Mixed-precision refinement of a block of right-hand sides: X0 = M^-1 B,
// then plain refinement steps R = B - Z X (one ZGEMM on the resident
// complex128 Z), X += M^-1 R on the columns whose backward error is above
// tol and still falling at least tenfold per step.  Columns that stall are
// finished one by one by gmres_ir.  Backward errors per column as in
// gmres_ir; be_hist records the largest over the columns.
static int mixed_solve_batch(mom_system *S, int64_t nrhs, double *B, int64_t ldb, double tol, int maxit,
                             int32_t *iters, double *berr) {
  hipStream_t s = S->ctx->stream;
  const int64_t n = S->n;
  const Lane L = main_lane(S->ctx);
  DevBuf dB, dX, dR, dT, dT2, dW, dW2, dT32, dperm, dn, dkeep;
  int rc;
  const size_t blk = sizeof(double2) * n * nrhs;
  if ((rc = dB.alloc(blk)) || (rc = dX.alloc(blk)) || (rc = dR.alloc(blk)) || (rc = dT.alloc(blk)) ||
      (rc = dT2.alloc(blk)) || (rc = dW.alloc(sizeof(double2) * NB * nrhs)) ||
      (rc = dW2.alloc(sizeof(double2) * NB * nrhs)) || (rc = dT32.alloc(sizeof(float2) * n * nrhs)) ||
      (rc = upload_perm(S, dperm)) || (rc = dn.alloc(sizeof(double) * 2 * nrhs)) ||
      (rc = dkeep.alloc(sizeof(int) * nrhs)))
    return rc;
  MOM_TRY(copy_h2d_2d(S->ctx, dB.ptr, n * sizeof(double2), B, ldb * sizeof(double2), n * sizeof(double2), nrhs));
  const dim3 g2((unsigned)((n + 255) / 256), (unsigned)nrhs);
  S->n_minv = S->n_matvec = 0;
  for (int i = 0; i < kBeHistory; i++) S->be_hist[i] = -1;
  // X (= or +=) C U^-1 L^-1 P R src, the solves in complex64
  auto minv = [&](const double2 *src, bool add) -> int {
    S->n_minv++;
    hipLaunchKernelGGL(scale_gather_round_kernel, g2, dim3(256), 0, s, src, n, S->rscale.as<double>(),
                       dperm.as<int>(), n, dT32.as<float2>());
    MOM_TRY(getrs_blocked32(S->ctx, S->Z32.as<float2>(), n, n, S->Linv.as<double2>(), S->Uinv.as<double2>(),
                            dT32.as<float2>(), nrhs, dW.as<double2>(), dW2.as<double2>()));
    hipLaunchKernelGGL(widen_scale_store_kernel, g2, dim3(256), 0, s, dX.as<double2>(), n, S->cscale.as<double>(),
                       dT32.as<float2>(), n, add ? 1 : 0);
    MOM_HIP(hipGetLastError());
    return 0;
  };
  auto colmax = [&](const double2 *v, std::vector<double> &out) -> int {
    hipLaunchKernelGGL(col_absmax_kernel, dim3((unsigned)nrhs), dim3(256), 0, s, v, n, n, dn.as<double>());
    MOM_HIP(hipGetLastError());
    out.resize(nrhs);
    return copy_d2h(S->ctx, out.data(), dn.ptr, sizeof(double) * nrhs);
  };
  std::vector<double> bn, rn, xn, be(nrhs, 0.0), prev(nrhs, 0.0);
  MOM_TRY(colmax(dB.as<double2>(), bn));
  MOM_TRY(minv(dB.as<double2>(), false));
  int it = 0;
  std::vector<int> keep(nrhs);
  // Stalled columns are finished one by one by GMRES, unless that would
  // cost more than factoring the resident complex128 matrix -- about 15
  // GMRES steps per column, each one application of the factors and one
  // product, against 8n^3/3 flops at about 60 TFLOP/s -- in which case the
  // system moves to FP64 factors and solves every column directly.  Once
  // the columns stalled so far settle that, the refinement stops at once.
  const double nd = (double)n;
  const double t_lu64 = 8.0 / 3.0 * nd * nd * nd / 6.0e13;
  const double t_col = 15.0 * (24.0 * nd * nd / 4.0e12 + 1.0e-3);
  auto fallback_settled = [&](int64_t nstall) { return g_force_fp64_fallback || (double)nstall * t_col > t_lu64; };
  for (;; it++) {
    S->n_matvec++;
    MOM_HIP(hipMemcpyAsync(dR.ptr, dB.ptr, blk, hipMemcpyDeviceToDevice, s));
    MOM_TRY(zgemm(L, n, nrhs, n, -1.0, false, S->Z64.as<double2>(), n, dX.as<double2>(), n, dR.as<double2>(), n));
    MOM_TRY(colmax(dR.as<double2>(), rn));
    MOM_TRY(colmax(dX.as<double2>(), xn));
    double bmax = 0;
    int nkeep = 0;
    for (int64_t j = 0; j < nrhs; j++) {
      be[j] = rn[j] / (S->znorm_inf * xn[j] + bn[j]);
      bmax = std::max(bmax, be[j]);
      keep[j] = be[j] > tol && (it == 0 || be[j] <= 0.1 * prev[j]) && !g_force_fp64_fallback;
      nkeep += keep[j];
    }
    if (it < kBeHistory) S->be_hist[it] = bmax;
    if (bmax <= tol || it >= maxit || nkeep == 0) break;
    int64_t nstall_now = 0;
    for (int64_t j = 0; j < nrhs; j++) nstall_now += be[j] > tol && !keep[j];
    if (nstall_now > 0 && fallback_settled(nstall_now)) break;
    MOM_TRY(copy_h2d(S->ctx, dkeep.ptr, keep.data(), sizeof(int) * nrhs));
    hipLaunchKernelGGL(mask_cols_kernel, g2, dim3(256), 0, s, dR.as<double2>(), n, n, dkeep.as<int>());
    MOM_TRY(minv(dR.as<double2>(), true));
    prev = be;
  }
  MOM_TRY(copy_d2h_2d(S->ctx, B, ldb * sizeof(double2), dX.ptr, n * sizeof(double2), n * sizeof(double2), nrhs));
  // Columns that stalled (typically near a resonance, where the FP32 factors'
  // backward error is large): GMRES one at a time, or FP64 factors for all
  // (until the next fill), as fallback_settled decides.
  int64_t nstall = 0;
  for (int64_t j = 0; j < nrhs; j++) nstall += be[j] > tol;
  if (nstall > 0 && fallback_settled(nstall)) {
    // the resident complex128 matrix is factored in place: if that fails
    // part-way it holds neither Z nor its factors, so the system needs a
    // new fill before it can factor or solve again
    int frc = factor_impl<double>(S, S->Z64.as<double2>());
    if (frc || S->info > 0) {
      S->filled = S->factored = false;
      return frc ? frc : fail(MOM_ERR_STATE, "the FP64 factorisation has a zero pivot");
    }
    S->precision = MOM_PREC_FP64;
    S->fell_back = true;
    DevBuf dperm64;
    MOM_TRY(upload_perm(S, dperm64));
    MOM_TRY(getrs_blocked<double>(S->ctx, S->Z64.as<double2>(), n, n, dperm64.as<int>(), S->Linv.as<double2>(),
                                  S->Uinv.as<double2>(), dB.as<double2>(), n, nrhs, dT2.as<double2>(),
                                  dW.as<double2>(), nullptr));
    MOM_TRY(copy_d2h_2d(S->ctx, B, ldb * sizeof(double2), dB.ptr, n * sizeof(double2), n * sizeof(double2), nrhs));
    if (iters) *iters = it;
    if (berr) *berr = -1.0;
    return 0;
  }
  const int nm = S->n_minv, nv = S->n_matvec;
  double hist[kBeHistory];
  std::copy(S->be_hist, S->be_hist + kBeHistory, hist);
  IrWork W;
  bool have_w = false;
  int itmax = it, extra_m = 0, extra_v = 0;
  double bemax = 0;
  for (int64_t j = 0; j < nrhs; j++) {
    if (be[j] <= tol) {
      bemax = std::max(bemax, be[j]);
      continue;
    }
    if (!have_w && (rc = ir_alloc(S, W))) return rc;
    have_w = true;
    std::vector<cd> b(n), x;
    MOM_TRY(copy_d2h(S->ctx, b.data(), dB.as<double2>() + j * n, sizeof(double2) * n));
    int itj = 0;
    double bej = 0;
    rc = gmres_ir(S, W, b, x, tol, maxit, &itj, &bej);
    if (rc && rc != MOM_ERR_NOCONV) return rc;
    extra_m += S->n_minv;
    extra_v += S->n_matvec;
    itmax = std::max(itmax, it + itj);
    bemax = std::max(bemax, bej);
    std::memcpy(B + 2 * j * ldb, x.data(), sizeof(double2) * n);
  }
  S->n_minv = nm + extra_m;
  S->n_matvec = nv + extra_v;
  std::copy(hist, hist + kBeHistory, S->be_hist);
  if (iters) *iters = itmax;
  if (berr) *berr = bemax;
  return bemax <= tol ? 0 : fail(MOM_ERR_NOCONV, "GMRES-IR did not reach the backward-error tolerance");
}

static int ir_alloc(mom_system *S, IrWork &W) {
  const int64_t n = S->n;
  int rc;
  if ((rc = upload_perm(S, W.dperm)) || (rc = W.v64.alloc(sizeof(double2) * n)) ||
      (rc = W.w64.alloc(sizeof(double2) * n)) || (rc = W.b64.alloc(sizeof(double2) * n)) ||
      (rc = W.x64.alloc(sizeof(double2) * n)) || (rc = W.t64.alloc(sizeof(double2) * n)) ||
      (rc = W.tmp64.alloc(sizeof(double2) * n)) || (rc = W.xb64.alloc(sizeof(double2) * NB)))
    return rc;
  return 0;
}

}  // namespace mom

using namespace mom;

extern "C" {

static int system_create(mom_ctx *ctx, int64_t nedges_total, int64_t order, int32_t precision, mom_system **sys) {
  if (!ctx || !sys) return fail(MOM_ERR_ARG, "NULL argument");
  if (precision != MOM_PREC_FP64 && precision != MOM_PREC_MIXED)
    return fail(MOM_ERR_ARG, "precision must be MOM_PREC_FP64 or MOM_PREC_MIXED");
  (void)hipSetDevice(ctx->device);
  mom_system *S = new mom_system;
  S->ctx = ctx;
  S->N = nedges_total;
  S->n = order;
  S->precision = precision;
  S->precision_req = precision;
  const int64_t n = S->n;
  int rc;
  if ((rc = S->Z64.alloc(sizeof(double2) * n * n)) || (rc = S->ipiv.alloc(sizeof(int) * n)) ||
      (rc = S->part.alloc(sizeof(double2) * n * kChunks)) || (rc = S->rscale.alloc(sizeof(double) * n)) ||
      (rc = S->cscale.alloc(sizeof(double) * n)) || (rc = S->rowsum.alloc(sizeof(double) * n)) ||
      (precision == MOM_PREC_MIXED && (rc = S->Z32.alloc(sizeof(float2) * n * n)))) {
    delete S;
    return rc;
  }
  *sys = S;
  return 0;
}

int mom_system_create(mom_ctx *ctx, int64_t nedges_total, int32_t precision, mom_system **sys) {
  if (nedges_total < 1 || nedges_total > (1 << 29)) return fail(MOM_ERR_ARG, "nedges_total out of range");
  return system_create(ctx, nedges_total, 2 * nedges_total, precision, sys);
}

int mom_system_create_order(mom_ctx *ctx, int64_t order, int32_t precision, mom_system **sys) {
  if (order < 2 || order > (1 << 30) || (order & 1)) return fail(MOM_ERR_ARG, "order must be even and in 2..2^30");
  return system_create(ctx, 0, order, precision, sys);
}

int mom_system_destroy(mom_system *sys) {
  if (!sys) return 0;
  (void)hipSetDevice(sys->ctx->device);
  (void)hipStreamSynchronize(sys->ctx->stream);
  delete sys;
  return 0;
}

int mom_system_set_pivoting(mom_system *sys, int32_t mode) {
  if (!sys) return fail(MOM_ERR_ARG, "NULL argument");
  if (mode != 0 && mode != 1) return fail(MOM_ERR_ARG, "pivoting must be 0 (partial) or 1 (tournament)");
  sys->pivoting = mode;
  return 0;
}

int mom_system_set_precision(mom_system *sys, int32_t precision) {
  if (!sys) return fail(MOM_ERR_ARG, "NULL argument");
  if (precision != MOM_PREC_FP64 && precision != MOM_PREC_MIXED) return fail(MOM_ERR_ARG, "unknown precision");
  if (precision == MOM_PREC_MIXED && !sys->Z32.ptr)
    return fail(MOM_ERR_ARG, "a system created MOM_PREC_FP64 has no complex64 copy");
  sys->precision_req = precision;
  if (sys->filled && !sys->factored) sys->precision = precision;
  return 0;
}

int mom_system_set_cooperative(mom_system *sys, int32_t allow) {
  if (!sys) return fail(MOM_ERR_ARG, "NULL argument");
  if (allow != 0 && allow != 1) return fail(MOM_ERR_ARG, "cooperative switch must be 0 or 1");
  sys->coop_off = allow == 0;
  return 0;
}

int mom_system_profile(mom_system *sys, int32_t enable) {
  if (!sys) return fail(MOM_ERR_ARG, "NULL argument");
  if (enable < 0 || enable > 2) return fail(MOM_ERR_ARG, "profile level must be 0, 1 or 2");
  sys->clock.enabled = enable != 0;
  sys->clock.all_phases = enable == 2;
  return 0;
}

int mom_system_fill(mom_system *S, int32_t ndom, const mom_mesh *const *doms, const double *ri, double omega,
                    int32_t nq, const double *qw, const double *qnodes) {
  if (!S) return fail(MOM_ERR_ARG, "NULL system");
  if (S->N == 0) return fail(MOM_ERR_STATE, "a system of a symmetry representation is filled by mom_system_fill_symm");
  (void)hipSetDevice(S->ctx->device);
  const double t0 = now_s();
  int rc = fill_system(S->ctx, ndom, doms, ri, S->N, omega, nq, qw, qnodes, S->Z64.as<double2>(), S->n);
  if (rc) return rc;
  MOM_HIP(hipStreamSynchronize(S->ctx->stream));
  S->t_fill = now_s() - t0;
  S->filled = true;
  S->znorm_one = 0;
  S->precision = S->precision_req;
  S->fell_back = false;
  S->factored = false;
  return 0;
}

int mom_system_fill_nls(mom_system *S, int32_t ndom, const mom_mesh *const *doms, const double *shri, double omega,
                        int32_t nq, const double *qw, const double *qnodes, int64_t nsrc, const double *src_coef,
                        int64_t ldsc, double *src, int64_t ldsrc) {
  if (!S || !doms || !shri) return fail(MOM_ERR_ARG, "NULL argument");
  if (S->N == 0) return fail(MOM_ERR_STATE, "the SH fill needs a system of the full problem");
  (void)hipSetDevice(S->ctx->device);
  const double t0 = now_s();
  int rc = fill_system_nls(S->ctx, ndom, doms, shri, S->N, omega, nq, qw, qnodes, S->Z64.as<double2>(), S->n, nsrc,
                           src_coef, ldsc, src, ldsrc);
  if (rc) return rc;
  MOM_HIP(hipStreamSynchronize(S->ctx->stream));
  S->t_fill = now_s() - t0;
  S->filled = true;
  S->znorm_one = 0;
  S->precision = S->precision_req;
  S->fell_back = false;
  S->factored = false;
  return 0;
}

int mom_system_fill_symm(int32_t nrep, mom_system *const *systems, int32_t ndom, const mom_mesh *const *doms,
                         const double *ri, int64_t nedges_total, const int32_t *bnd, int32_t ngen,
                         const int32_t *planes, double omega, int32_t nq, const double *qw, const double *qnodes) {
  if (!systems || !doms || !ri || !bnd) return fail(MOM_ERR_ARG, "NULL argument");
  if (nedges_total < 1 || ndom < 1) return fail(MOM_ERR_ARG, "bad sizes");
  Group G;
  MOM_TRY(symm_group(ngen, planes, &G));
  if (nrep != G.n) return fail(MOM_ERR_ARG, "one system per representation of the group is needed");
  double2 *Z[kMaxReps];
  int64_t ld[kMaxReps], order[kMaxReps];
  for (int r = 0; r < nrep; r++) {
    mom_system *S = systems[r];
    if (!S) return fail(MOM_ERR_ARG, "NULL system");
    if (S->ctx != systems[0]->ctx) return fail(MOM_ERR_ARG, "systems belong to different contexts");
    if (S->N != 0) return fail(MOM_ERR_ARG, "representation systems are made by mom_system_create_order");
    Z[r] = S->Z64.as<double2>();
    ld[r] = S->n;
    order[r] = S->n;
  }
  mom_ctx *ctx = systems[0]->ctx;
  (void)hipSetDevice(ctx->device);
  const double t0 = now_s();
  MOM_TRY(fill_system_symm(ctx, ndom, doms, ri, nedges_total, bnd, G, omega, nq, qw, qnodes, Z, ld, order));
  const double t = now_s() - t0;
  for (int r = 0; r < nrep; r++) {
    mom_system *S = systems[r];
    S->t_fill = t;
    S->filled = true;
    S->znorm_one = 0;
    S->precision = S->precision_req;
    S->fell_back = false;
    S->factored = false;
  }
  return 0;
}

int mom_system_get_matrix(mom_system *S, double *A, int64_t lda) {
  if (!S || !A) return fail(MOM_ERR_ARG, "NULL argument");
  if (lda < S->n) return fail(MOM_ERR_ARG, "lda < 2N");
  if (!S->filled || (S->factored && S->precision == MOM_PREC_FP64))
    return fail(MOM_ERR_STATE, "no unfactored matrix on the device");
  const int64_t n = S->n;
  MOM_TRY(copy_d2h_2d(S->ctx, A, lda * sizeof(double2), S->Z64.ptr, n * sizeof(double2), n * sizeof(double2), n));
  return 0;
}

int mom_system_set_matrix(mom_system *S, const double *A, int64_t lda) {
  if (!S || !A) return fail(MOM_ERR_ARG, "NULL argument");
  if (lda < S->n) return fail(MOM_ERR_ARG, "lda < 2N");
  const int64_t n = S->n;
  MOM_TRY(copy_h2d_2d(S->ctx, S->Z64.ptr, n * sizeof(double2), A, lda * sizeof(double2), n * sizeof(double2), n));
  MOM_HIP(hipStreamSynchronize(S->ctx->stream));
  S->filled = true;
  S->znorm_one = 0;
  S->precision = S->precision_req;
  S->fell_back = false;
  S->factored = false;
  return 0;
}

int mom_system_factor(mom_system *S) {
  if (!S) return fail(MOM_ERR_ARG, "NULL system");
  if (!S->filled) return fail(MOM_ERR_STATE, "factor before fill");
  if (S->factored) return fail(MOM_ERR_STATE, "already factored");
  (void)hipSetDevice(S->ctx->device);
  hipStream_t s = S->ctx->stream;
  const int64_t n = S->n;
  const double t0 = now_s();
  int rc;
  if (S->precision == MOM_PREC_FP64) {
    rc = factor_impl<double>(S, S->Z64.as<double2>());
  } else {
    hipLaunchKernelGGL(row_stats_part_kernel, dim3((unsigned)((n + 255) / 256), kChunks), dim3(256), 0, s,
                       S->Z64.as<double2>(), n, n, S->part.as<double2>());
    hipLaunchKernelGGL(row_stats_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       S->part.as<double2>(), n, S->rscale.as<double>(), S->rowsum.as<double>());
    hipLaunchKernelGGL(col_scale_kernel, dim3((unsigned)n), dim3(256), 0, s, S->Z64.as<double2>(), n, n,
                       S->rscale.as<double>(), S->cscale.as<double>());
    hipLaunchKernelGGL(to_fp32_kernel, dim3((unsigned)std::min<int64_t>(64, (n + 255) / 256), (unsigned)n),
                       dim3(256), 0, s, S->Z64.as<double2>(), n, n, S->rscale.as<double>(), S->cscale.as<double>(),
                       S->Z32.as<float2>());
    MOM_HIP(hipGetLastError());
    std::vector<double> rs(n);
    MOM_TRY(copy_d2h(S->ctx, rs.data(), S->rowsum.ptr, sizeof(double) * n));
    S->znorm_inf = *std::max_element(rs.begin(), rs.end());
    rc = factor_impl<float>(S, S->Z32.as<float2>());
  }
  if (rc) return rc;
  S->t_factor = now_s() - t0;
  S->factored = true;
  return S->info;
}

int mom_system_solve(mom_system *S, int64_t nrhs, double *B, int64_t ldb, double tol, int32_t maxit,
                     int32_t *iters, double *berr) {
  if (!S || !B) return fail(MOM_ERR_ARG, "NULL argument");
  if (!S->factored) return fail(MOM_ERR_STATE, "solve before factor");
  if (nrhs < 1 || ldb < S->n) return fail(MOM_ERR_ARG, "bad sizes");
  (void)hipSetDevice(S->ctx->device);
  hipStream_t s = S->ctx->stream;
  const int64_t n = S->n;
  const double t0 = now_s();
  int rc;
  if (S->precision == MOM_PREC_FP64) {
    if (S->info > 0) return fail(MOM_ERR_STATE, "the factorisation has a zero pivot");
    DevBuf dB, tmp, xb, dperm;
    if ((rc = dB.alloc(sizeof(double2) * n * nrhs)) || (rc = tmp.alloc(sizeof(double2) * n * nrhs)) ||
        (rc = xb.alloc(sizeof(double2) * NB * nrhs)) || (rc = upload_perm(S, dperm)))
      return rc;
    MOM_TRY(copy_h2d_2d(S->ctx, dB.ptr, n * sizeof(double2), B, ldb * sizeof(double2), n * sizeof(double2), nrhs));
    if (nrhs >= kBatchRhs)
      rc = getrs_blocked<double>(S->ctx, S->Z64.as<double2>(), n, n, dperm.as<int>(), S->Linv.as<double2>(),
                                 S->Uinv.as<double2>(), dB.as<double2>(), n, nrhs, tmp.as<double2>(),
                                 xb.as<double2>(), nullptr);
    else
      rc = getrs<double, double>(s, S->Z64.as<double2>(), n, n, dperm.as<int>(), S->Linv.as<double2>(),
                                 S->Uinv.as<double2>(), dB.as<double2>(), n, nrhs, tmp.as<double2>(), xb.as<double2>());
    if (rc) return rc;
    MOM_TRY(copy_d2h_2d(S->ctx, B, ldb * sizeof(double2), dB.ptr, n * sizeof(double2), n * sizeof(double2), nrhs));
    if (iters) *iters = 0;
    if (berr) *berr = -1.0;
  } else if (nrhs >= kBatchRhs) {
    if ((rc = mixed_solve_batch(S, nrhs, B, ldb, tol, maxit, iters, berr))) return rc;
  } else {
    IrWork W;
    if ((rc = ir_alloc(S, W))) return rc;
    int itmax = 0;
    double bemax = 0;
    for (int64_t r = 0; r < nrhs; r++) {
      std::vector<cd> b(n), x;
      std::memcpy(b.data(), B + 2 * r * ldb, sizeof(double2) * n);
      int it = 0;
      double be = 0;
      rc = gmres_ir(S, W, b, x, tol, maxit, &it, &be);
      itmax = std::max(itmax, it);
      bemax = std::max(bemax, be);
      std::memcpy(B + 2 * r * ldb, x.data(), sizeof(double2) * n);
      if (rc && rc != MOM_ERR_NOCONV) return rc;
    }
    if (iters) *iters = itmax;
    if (berr) *berr = bemax;
    if (bemax > tol) return fail(MOM_ERR_NOCONV, "GMRES-IR did not reach the backward-error tolerance");
  }
  S->t_solve = now_s() - t0;
  return 0;
}

int mom_system_solve_adjoint(mom_system *S, int64_t nrhs, double *B, int64_t ldb, double tol, int32_t maxit,
                             int32_t *iters, double *berr) {
  if (!S || !B) return fail(MOM_ERR_ARG, "NULL argument");
  if (!S->factored) return fail(MOM_ERR_STATE, "solve before factor");
  if (S->info > 0) return fail(MOM_ERR_STATE, "the factorisation has a zero pivot");
  if (nrhs < 1 || ldb < S->n) return fail(MOM_ERR_ARG, "bad sizes");
  (void)hipSetDevice(S->ctx->device);
  hipStream_t s = S->ctx->stream;
  const int64_t n = S->n;
  const double t0 = now_s();
  int rc;
  if (S->precision == MOM_PREC_FP64) {
    const int64_t chunk = std::min<int64_t>(nrhs, 4096);
    DevBuf dB, tmp, xb, dperm;
    if ((rc = dB.alloc(sizeof(double2) * n * chunk)) || (rc = tmp.alloc(sizeof(double2) * n * chunk)) ||
        (rc = xb.alloc(sizeof(double2) * NB * chunk)) || (rc = upload_perm(S, dperm)))
      return rc;
    for (int64_t c0 = 0; c0 < nrhs; c0 += chunk) {
      const int64_t nc = std::min(chunk, nrhs - c0);
      double *Bc = B + 2 * c0 * ldb;
      MOM_TRY(copy_h2d_2d(S->ctx, dB.ptr, n * sizeof(double2), Bc, ldb * sizeof(double2), n * sizeof(double2), nc));
      MOM_TRY((getrs_h<double, double>(s, S->Z64.as<double2>(), n, n, dperm.as<int>(), S->Linv.as<double2>(),
                                       S->Uinv.as<double2>(), dB.as<double2>(), n, nc, tmp.as<double2>(),
                                       xb.as<double2>())));
      MOM_TRY(copy_d2h_2d(S->ctx, Bc, ldb * sizeof(double2), dB.ptr, n * sizeof(double2), n * sizeof(double2), nc));
    }
    if (iters) *iters = 0;
    if (berr) *berr = -1.0;
  } else {
    if (S->znorm_one == 0) {
      DevBuf cs;
      MOM_TRY(cs.alloc(sizeof(double) * n));
      hipLaunchKernelGGL(col_abssum_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, S->Z64.as<double2>(), n,
                         n, cs.as<double>());
      MOM_HIP(hipGetLastError());
      std::vector<double> h(n);
      MOM_TRY(copy_d2h(S->ctx, h.data(), cs.ptr, sizeof(double) * n));
      S->znorm_one = *std::max_element(h.begin(), h.end());
    }
    IrWork W;
    if ((rc = ir_alloc(S, W))) return rc;
    W.adj = true;
    int itmax = 0;
    double bemax = 0;
    for (int64_t r = 0; r < nrhs; r++) {
      std::vector<cd> b(n), x;
      std::memcpy(b.data(), B + 2 * r * ldb, sizeof(double2) * n);
      int it = 0;
      double be = 0;
      rc = gmres_ir(S, W, b, x, tol, maxit, &it, &be);
      itmax = std::max(itmax, it);
      bemax = std::max(bemax, be);
      std::memcpy(B + 2 * r * ldb, x.data(), sizeof(double2) * n);
      if (rc && rc != MOM_ERR_NOCONV) return rc;
    }
    if (iters) *iters = itmax;
    if (berr) *berr = bemax;
    if (bemax > tol) return fail(MOM_ERR_NOCONV, "GMRES-IR did not reach the backward-error tolerance");
  }
  S->t_solve = now_s() - t0;
  return 0;
}

int mom_system_solve_pw(mom_system *S, const mom_mesh *mesh, double omega, double ri_re, double ri_im, int64_t nsrc,
                        const double *angles, int32_t nq, const double *qw, const double *qnodes, double tol,
                        int32_t maxit, double *X, int64_t ldx, int32_t *iters, double *berr) {
  if (!S || !mesh || !X) return fail(MOM_ERR_ARG, "NULL argument");
  if (S->N == 0) return fail(MOM_ERR_STATE, "a system of a symmetry representation takes reduced sources");
  if (ldx < S->n || nsrc < 1) return fail(MOM_ERR_ARG, "bad sizes");
  int rc = mom_srcvec_pw(S->ctx, mesh, S->N, omega, ri_re, ri_im, nsrc, angles, nq, qw, qnodes, X, ldx);
  if (rc) return rc;
  return mom_system_solve(S, nsrc, X, ldx, tol, maxit, iters, berr);
}

int mom_system_stats(const mom_system *S, double *out) {
  if (!S || !out) return fail(MOM_ERR_ARG, "NULL argument");
  out[0] = S->t_fill;
  out[1] = S->t_factor;
  out[2] = S->t_solve;
  out[3] = S->clock.seconds[mom::kPhGemm];
  out[4] = S->clock.launches[mom::kPhGemm];
  out[5] = S->clock.flops[mom::kPhGemm];
  out[6] = S->clock.seconds[mom::kPhPanel];
  out[7] = S->clock.seconds[mom::kPhTrsm];
  out[8] = S->clock.seconds[mom::kPhLaswp];
  out[9] = S->n_minv;
  out[10] = S->n_matvec;
  for (int i = 0; i < mom::kBeHistory; i++) out[11 + i] = S->be_hist[i];
  out[19] = S->clock.seconds[mom::kPhGemmNarrow];
  out[20] = S->clock.launches[mom::kPhGemmNarrow];
  out[21] = S->clock.flops[mom::kPhGemmNarrow];
  out[22] = S->fell_back ? 1.0 : 0.0;
  return 0;
}

int mom_solve_multi_linsys(mom_ctx *ctx, int64_t n, double *A, int64_t lda, int64_t nrhs, double *B, int64_t ldb,
                           int32_t *ipiv_out) {
  if (!ctx || !A || !B) return fail(MOM_ERR_ARG, "NULL argument");
  if (n < 1 || lda < n || ldb < n || nrhs < 1) return fail(MOM_ERR_ARG, "bad sizes");
  (void)hipSetDevice(ctx->device);
  hipStream_t s = ctx->stream;
  mom_system S;
  S.ctx = ctx;
  S.n = n;
  S.N = (n + 1) / 2;
  int rc;
  if ((rc = S.Z64.alloc(sizeof(double2) * n * n)) || (rc = S.ipiv.alloc(sizeof(int) * n))) return rc;
  MOM_TRY(copy_h2d_2d(ctx, S.Z64.ptr, n * sizeof(double2), A, lda * sizeof(double2), n * sizeof(double2), n));
  if ((rc = factor_impl<double>(&S, S.Z64.as<double2>()))) return rc;
  S.factored = true;
  S.precision = MOM_PREC_FP64;
  int info = S.info;
  DevBuf dB, tmp, xb, dperm;
  if ((rc = dB.alloc(sizeof(double2) * n * nrhs)) || (rc = tmp.alloc(sizeof(double2) * n * nrhs)) ||
      (rc = xb.alloc(sizeof(double2) * NB * nrhs)) || (rc = upload_perm(&S, dperm)))
    return rc;
  if (info == 0) {
    MOM_TRY(copy_h2d_2d(ctx, dB.ptr, n * sizeof(double2), B, ldb * sizeof(double2), n * sizeof(double2), nrhs));
    if ((rc = getrs<double, double>(s, S.Z64.as<double2>(), n, n, dperm.as<int>(), S.Linv.as<double2>(), S.Uinv.as<double2>(),
                            dB.as<double2>(), n, nrhs, tmp.as<double2>(), xb.as<double2>())))
      return rc;
    MOM_TRY(copy_d2h_2d(ctx, B, ldb * sizeof(double2), dB.ptr, n * sizeof(double2), n * sizeof(double2), nrhs));
  }
  MOM_TRY(copy_d2h_2d(ctx, A, lda * sizeof(double2), S.Z64.ptr, n * sizeof(double2), n * sizeof(double2), n));
  if (ipiv_out) MOM_TRY(copy_d2h(ctx, ipiv_out, S.ipiv.ptr, sizeof(int) * n));
  return info;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Test hooks (include/mom_gpu_testing.h).
namespace mom {
template <typename T>
__global__ void mfma_probe_kernel(double *out) {
  using acc4 = typename Cx<T>::acc4;
  const int L0 = blockIdx.x / 64, L1 = blockIdx.x % 64, lane = threadIdx.x;
  const T a = lane == L0 ? (T)1 : (T)0, b = lane == L1 ? (T)1 : (T)0;
  acc4 c;
  for (int v = 0; v < 4; v++) c[v] = 0;
  c = mfma4<T>(a, b, c);
  for (int v = 0; v < 4; v++) out[((int64_t)blockIdx.x * 64 + lane) * 4 + v] = (double)c[v];
}

__global__ void widen_kernel(const float2 *in, double2 *out, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_double2(in[i].x, in[i].y);
}
__global__ void narrow_kernel(const double2 *in, float2 *out, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_float2((float)in[i].x, (float)in[i].y);
}
}  // namespace mom

extern "C" int mom_test_mfma_layout(mom_ctx *ctx, int32_t precision, double *out) {
  if (!ctx || !out) return fail(MOM_ERR_ARG, "NULL argument");
  DevBuf d;
  const size_t cnt = 64 * 64 * 64 * 4;
  int rc = d.alloc(cnt * sizeof(double));
  if (rc) return rc;
  if (precision == 64)
    hipLaunchKernelGGL(mfma_probe_kernel<double>, dim3(4096), dim3(64), 0, ctx->stream, d.as<double>());
  else
    hipLaunchKernelGGL(mfma_probe_kernel<float>, dim3(4096), dim3(64), 0, ctx->stream, d.as<double>());
  MOM_HIP(hipGetLastError());
  MOM_TRY(copy_d2h(ctx, out, d.ptr, cnt * sizeof(double)));
  return 0;
}

extern "C" int mom_test_panel_mode(int32_t mode) {
  if (mode != 0 && mode != 1) return fail(MOM_ERR_ARG, "panel mode must be 0 or 1");
  g_panel_mode = mode;
  return 0;
}

extern "C" int mom_test_u_solve(int32_t mode) {
  if (mode != 0 && mode != 1) return fail(MOM_ERR_ARG, "U-block mode must be 0 or 1");
  g_u_subst = mode;
  return 0;
}

extern "C" int mom_test_tournament(int32_t enable) {
  g_tournament = enable ? 1 : 0;
  return 0;
}

extern "C" int mom_test_batch_fallback(int32_t force) {
  g_force_fp64_fallback = force ? 1 : 0;
  return 0;
}

extern "C" int mom_test_factor_graph(int32_t enable, int64_t *counts) {
  if (enable != 0 && enable != 1 && enable != -1) return fail(MOM_ERR_ARG, "graph switch must be -1, 0 or 1");
  if (enable >= 0) g_factor_graph = enable;
  if (counts) {
    counts[0] = g_graph_counts[0];
    counts[1] = g_graph_counts[1];
  }
  return 0;
}

extern "C" int mom_test_lookahead(int32_t enable) {
  if (enable != 0 && enable != 1) return fail(MOM_ERR_ARG, "look-ahead switch must be 0 or 1");
  g_lookahead = enable;
  return 0;
}

extern "C" int mom_test_unit_lower_inverse(mom_ctx *ctx, const double *L, int64_t ldl, double *E) {
  if (!ctx || !L || !E) return fail(MOM_ERR_ARG, "NULL argument");
  if (ldl < NB) return fail(MOM_ERR_ARG, "bad leading dimension");
  hipStream_t s = ctx->stream;
  DevBuf dL, fL, fE, dE;
  int rc;
  if ((rc = dL.alloc(sizeof(double2) * ldl * NB)) || (rc = fL.alloc(sizeof(float2) * ldl * NB)) ||
      (rc = fE.alloc(sizeof(float2) * NB * NB)) || (rc = dE.alloc(sizeof(double2) * NB * NB)))
    return rc;
  MOM_TRY(copy_h2d(ctx, dL.ptr, L, sizeof(double2) * ldl * NB));
  hipLaunchKernelGGL(narrow_kernel, dim3((unsigned)((ldl * NB + 255) / 256)), dim3(256), 0, s, dL.as<double2>(),
                     fL.as<float2>(), ldl * NB);
  hipLaunchKernelGGL(unit_lower_inverse_kernel, dim3(NB / kInvCols), dim3(256), 0, s, fL.as<float2>(), ldl, fE.as<float2>());
  hipLaunchKernelGGL(widen_kernel, dim3((unsigned)((NB * NB + 255) / 256)), dim3(256), 0, s, fE.as<float2>(),
                     dE.as<double2>(), (int64_t)NB * NB);
  MOM_HIP(hipGetLastError());
  return copy_d2h(ctx, E, dE.ptr, sizeof(double2) * NB * NB);
}

extern "C" int mom_test_gemm_sub(mom_ctx *ctx, int32_t precision, int64_t M, int64_t N, int64_t K, const double *A,
                                 int64_t lda, const double *B, int64_t ldb, double *C, int64_t ldc) {
  if (!ctx || !A || !B || !C) return fail(MOM_ERR_ARG, "NULL argument");
  if (M < 1 || N < 1 || K < 1 || lda < M || ldb < K || ldc < M) return fail(MOM_ERR_ARG, "bad sizes");
  hipStream_t s = ctx->stream;
  DevBuf dA, dB, dC;
  int rc;
  if ((rc = dA.alloc(sizeof(double2) * lda * K)) || (rc = dB.alloc(sizeof(double2) * ldb * N)) ||
      (rc = dC.alloc(sizeof(double2) * ldc * N)))
    return rc;
  MOM_TRY(copy_h2d(ctx, dA.ptr, A, sizeof(double2) * lda * K));
  MOM_TRY(copy_h2d(ctx, dB.ptr, B, sizeof(double2) * ldb * N));
  MOM_TRY(copy_h2d(ctx, dC.ptr, C, sizeof(double2) * ldc * N));
  if (precision == 64) {
    if ((rc = gemm_sub<double>(main_lane(ctx), dA.as<double2>(), lda, dB.as<double2>(), ldb, dC.as<double2>(), ldc, M, N, K)))
      return rc;
  } else {
    DevBuf fA, fB, fC;
    if ((rc = fA.alloc(sizeof(float2) * lda * K)) || (rc = fB.alloc(sizeof(float2) * ldb * N)) ||
        (rc = fC.alloc(sizeof(float2) * ldc * N)))
      return rc;
    auto nar = [&](DevBuf &in, DevBuf &out, int64_t cnt) {
      hipLaunchKernelGGL(narrow_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, in.as<double2>(),
                         out.as<float2>(), cnt);
    };
    nar(dA, fA, lda * K);
    nar(dB, fB, ldb * N);
    nar(dC, fC, ldc * N);
    if ((rc = gemm_sub<float>(main_lane(ctx), fA.as<float2>(), lda, fB.as<float2>(), ldb, fC.as<float2>(), ldc, M, N, K)))
      return rc;
    hipLaunchKernelGGL(widen_kernel, dim3((unsigned)((ldc * N + 255) / 256)), dim3(256), 0, s, fC.as<float2>(),
                       dC.as<double2>(), ldc * N);
  }
  MOM_HIP(hipGetLastError());
  return copy_d2h(ctx, C, dC.ptr, sizeof(double2) * ldc * N);
}
