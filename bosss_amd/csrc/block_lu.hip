// Batched uneven cell-block LU, inverse and block-Jacobi sweep (BlockJacobi.cs:63-197,
// IMatrixExtensions.InvertInPlace:2128-2230 -> LAPACK DGETRF/DGETRI semantics).
//
// Every form follows DGETF2: partial pivoting picks the first row with the largest
// |a| in the column, as LAPACK's idamax does (a NaN in the diagonal position is never
// beaten); a zero pivot sets INFO = j+1 (1-based) and the factorisation continues.
// Pivots are stored 1-based (ipiv).  Blocks up to 20 DOF keep their rows in registers
// (several blocks per wavefront), 21..128 DOF one block per workgroup in registers,
// larger ones (Schwarz blocks) go through global memory.
//
// All blocks of a batch are processed by one launch (blocks of different size
// share the launch; each wave reads its own size), so a whole block-Jacobi set
// of uneven XDG blocks (10x10 and 20x20) is one kernel.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bosss_hip.h"
#include "devbuf.hpp"

namespace bosss {

constexpr int LU_MAXN = 128;

template <typename Real>
__device__ __forceinline__ void wave_argmax(Real& v, int& idx) {
  // reduce (|a|, index) over the 64 lanes; ties keep the smaller index
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    Real ov = __shfl_xor(v, o, 64);
    int oi = __shfl_xor(idx, o, 64);
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
}

// Shared state of the register-tile kernel.  The row placement (pos, lrow) is kept once per
// wavefront: each wavefront applies every exchange to its own copy, so a column step needs one
// workgroup barrier (publishing the column and the pivot) and no wavefront reads a copy that
// another one is still writing.
template <int RA>
struct TileLDS {
  double col[2][16 * RA], row[2][16 * RA], pv[2];
  int pos[4][16 * RA], lrow[4][16 * RA], piv[16 * RA], p[2];
};

// Loads block g (n x n, row-major) into the register tile and resets the row bookkeeping.
template <int RA>
__device__ __forceinline__ void tile_load(double (&a)[RA][RA], const double* __restrict__ g, int n, TileLDS<RA>& L) {
  const int tid = threadIdx.x, tc = tid >> 4, tr = tid & 15;
#pragma unroll
  for (int x = 0; x < RA; x++)
#pragma unroll
    for (int y = 0; y < RA; y++) {
      const int r = tr + 16 * x, c = tc + 16 * y;
      a[x][y] = 0.0;
      if (r < n && c < n) a[x][y] = g[(int64_t)r * n + c];
    }
  for (int i = tid; i < 4 * 16 * RA; i += 256) {
    const int w = i / (16 * RA), r = i - w * 16 * RA;
    L.pos[w][r] = r;
    L.lrow[w][r] = r;
  }
  __syncthreads();
}

// Column steps 0..n-1 of getf2 (GJ = false) or of in-place Gauss-Jordan (GJ = true) on the register
// tile.  Returns INFO (first zero pivot, 1-based; Gauss-Jordan stops there).  L.piv[j] receives the
// position of step j's pivot row, L.pos / L.lrow the final row placement (every wavefront's copy).
//
// One barrier per step: the 16 threads holding column j publish it and the pivot (row, value);
// after the barrier the thread of each 16-lane group that holds the pivot row's entries of the
// group's columns c = tc + 16 y publishes them in row[][c], which only that group reads back.
// Column group tc lives in wavefront tc / 4 and the LDS accesses of one wavefront complete in
// order, so this exchange needs no workgroup barrier; lane 0 of each wavefront applies the row
// exchange to that wavefront's copy of the placement.
template <int RA, bool GJ>
__device__ __forceinline__ int tile_eliminate(double (&a)[RA][RA], int n, TileLDS<RA>& L) {
  const int tid = threadIdx.x, tc = tid >> 4, tr = tid & 15, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  int* pos = L.pos[w];
  int* lrow = L.lrow[w];
  unsigned used = 0;   // bit x: row tr + 16 x is in a position <= j (LU form only)
  int inf = 0;
  for (int j = 0; j < n; j++) {
    const int k = j & 1, jc = j & 15, jy = j >> 4;
    if (tc == jc) {
      // the 16 threads holding column j: publish it, find the pivot among the unused rows
      double best = -1.0, bval = 0.0;
      int bpos = 1 << 30, bp = -1;
#pragma unroll
      for (int yy = 0; yy < RA; yy++)
        if (yy == jy) {
#pragma unroll
          for (int x = 0; x < RA; x++) {
            const int r = tr + 16 * x;
            const double v0 = a[x][yy];
            if (r < n) {
              L.col[k][r] = v0;
              const int ps = pos[r];
              if (ps >= j) {
                double v = fabs(v0);
                if (ps == j && v != v) v = (double)INFINITY;   // idamax: a NaN first element is never beaten
                if (v > best || (v == best && ps < bpos)) { best = v; bpos = ps; bp = r; bval = v0; }
              }
            }
            if (GJ) a[x][yy] = 0.0;   // becomes -a[r][j] / pivot through the update
          }
        }
#pragma unroll
      for (int o = 8; o >= 1; o >>= 1) {
        const double ov = __shfl_xor(best, o, 16), oval = __shfl_xor(bval, o, 16);
        const int opos = __shfl_xor(bpos, o, 16), op = __shfl_xor(bp, o, 16);
        if (ov > best || (ov == best && opos < bpos)) { best = ov; bpos = opos; bp = op; bval = oval; }
      }
      if (tr == 0) {
        if (bp < 0) { bp = lrow[j]; bval = L.col[k][bp]; }   // no comparable candidate (all NaN): position j
        L.p[k] = bp;
        L.pv[k] = bval;
      }
    }
    __syncthreads();
    const int p = __builtin_amdgcn_readfirstlane(L.p[k]);
    const double pv = L.pv[k];
    if (tid == 0) L.piv[j] = pos[p];
    if (!(pv != 0.0)) {   // zero pivot (NaN passes, as in DGETF2)
      if (!inf) inf = j + 1;
      if (GJ) return inf;
      const int r0 = lrow[j];   // no exchange: the row in position j is done all the same
      if (tr == (r0 & 15)) used |= 1u << (r0 >> 4);
      continue;
    }
    if (lane == 0) {   // read by this wavefront's lanes from the next step on
      const int q = pos[p], rj = lrow[j];
      pos[p] = j; pos[rj] = q; lrow[j] = p; lrow[q] = rj;
    }
    const int pc = p & 15, px = __builtin_amdgcn_readfirstlane(p >> 4);
    const double pinv = 1.0 / pv;
    if (tr == pc) {   // the pivot row's entries of the group's columns tc + 16 y
#pragma unroll
      for (int xx = 0; xx < RA; xx++)
        if (xx == px) {
#pragma unroll
          for (int y = 0; y < RA; y++) {
            const int c = tc + 16 * y;
            L.row[k][c] = GJ ? (c == j ? pinv : a[xx][y] * pinv) : a[xx][y];
          }
        }
    }
    __builtin_amdgcn_wave_barrier();
    double u[RA];
#pragma unroll
    for (int y = 0; y < RA; y++) u[y] = L.row[k][tc + 16 * y];
    if (GJ) {
#pragma unroll
      for (int x = 0; x < RA; x++) {
        const int r = tr + 16 * x;
        const double f = r < n ? L.col[k][r] : 0.0;
#pragma unroll
        for (int y = 0; y < RA; y++) a[x][y] -= f * u[y];
      }
      if (tr == pc) {
#pragma unroll
        for (int xx = 0; xx < RA; xx++)
          if (xx == px) {
#pragma unroll
            for (int y = 0; y < RA; y++) a[xx][y] = u[y];
          }
      }
    } else {
      // rows after the pivot in factor order: multiplier into column j, columns > j updated
#pragma unroll
      for (int x = 0; x < RA; x++) {
        const int r = tr + 16 * x;
        if (r < n && r != p && !((used >> x) & 1u)) {
          const double l = L.col[k][r] * pinv;
#pragma unroll
          for (int y = 0; y < RA; y++)
            if (tc + 16 * y > j) a[x][y] -= l * u[y];
          if (tc == jc) {
#pragma unroll
            for (int yy = 0; yy < RA; yy++)
              if (yy == jy) a[x][yy] = l;
          }
        }
      }
      if (tr == pc) used |= 1u << px;
    }
  }
  return inf;
}

// Writes the tile back: factors in position order (row pos[r]), or the inverse with the column
// permutation too (X[pos[r]][lrow[c]] = a[r][c]).
template <int RA>
__device__ __forceinline__ void tile_store(const double (&a)[RA][RA], double* __restrict__ g, int n, const TileLDS<RA>& L, bool inverse) {
  const int tid = threadIdx.x, tc = tid >> 4, tr = tid & 15;
#pragma unroll
  for (int x = 0; x < RA; x++) {
    const int r = tr + 16 * x;
    if (r < n) {
      const int64_t row = (int64_t)L.pos[0][r] * n;
#pragma unroll
      for (int y = 0; y < RA; y++) {
        const int c = tc + 16 * y;
        if (c < n) g[row + (inverse ? L.lrow[0][c] : c)] = a[x][y];
      }
    }
  }
}

// Waves per SIMD the register-tile kernel is compiled for.  The 100-DOF blocks (RA = 7) with the
// inverse need 186 VGPRs unconstrained (two waves); at three waves (168 VGPRs, an 88 B per-lane
// spill) the 262 144 blocks of p = 3 Euler factor and invert in 95 ms instead of 122 ms, the
// column steps being latency bound (four waves spill 240 B and take 329 ms).
template <int RA>
constexpr int lu_tile_waves() { return RA == 7 ? 3 : 1; }

// Blocks of 21..16 RA DOF (the 100-DOF cell blocks of p = 3 Euler, up to 128): one workgroup of
// 256 threads per block, the whole block in registers.  Thread (tc, tr) = (tid >> 4, tid & 15)
// holds rows tr + 16 x and columns tc + 16 y (x, y < RA), in the block's original row order:
// row exchanges are bookkeeping only, pos[r] is the position (row of the factors) that physical
// row r has been given and lrow[] its inverse.  Each column step: the 16 threads holding column j
// find the pivot among the rows not yet used (largest |a|, ties to the smaller position, a NaN in
// position j never beaten: getf2's IDAMAX in getf2's row order) with shuffles inside their 16-lane
// group and publish the column; the threads holding the pivot row publish it to their own
// 16-lane groups; every thread updates its entries.  Factors, pivots and INFO are getf2's, operation for operation (multiplier a * (1 / pivot),
// then the rank-one update of the columns after j).
//
// With INV the elimination is Gauss-Jordan in place (every row other than the pivot row updated at
// every step: uniform work for all threads) and the inverse is written out with the row and column
// permutation applied.  A zero pivot stops it; the block is then factored from its original values
// instead (LU factors and INFO, no inverse, as DGETRF leaves them when DGETRI cannot proceed).  Register columns and rows are
// picked with wavefront-uniform branches, so no register array is indexed dynamically (the block
// stays in registers; RA = 7 with the inverse spills 88 B per lane at the three waves above).
template <int RA, bool INV>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(lu_tile_waves<RA>()))) k_block_lu_tile(int nblk, int nmax, const int32_t* __restrict__ nsz, const int64_t* __restrict__ off,
                                                       double* __restrict__ A, int32_t* __restrict__ ipiv,
                                                       const int64_t* __restrict__ pivoff, int32_t* __restrict__ info) {
  __shared__ TileLDS<RA> L;
  const int b = blockIdx.x;
  if (b >= nblk) return;
  const int n = nsz[b];
  if (n <= 0 || n > nmax || n > 16 * RA) {   // size outside what the launch was sized for: report, touch nothing
    if (threadIdx.x == 0 && info) info[b] = -1;
    return;
  }
  double* g = A + off[b];
  double a[RA][RA];
  tile_load<RA>(a, g, n, L);
  int inf = 0;
  bool inverse = false;
  if (INV) {
    inf = tile_eliminate<RA, true>(a, n, L);
    inverse = inf == 0;
    if (!inverse) {   // singular: factor the original block (LU and INFO, no inverse)
      __syncthreads();
      tile_load<RA>(a, g, n, L);
    }
  }
  if (!inverse) inf = tile_eliminate<RA, false>(a, n, L);
  __syncthreads();
  tile_store<RA>(a, g, n, L, inverse);
  if (ipiv)
    for (int i = threadIdx.x; i < n; i += 256) ipiv[pivoff[b] + i] = L.piv[i] + 1;
  if (threadIdx.x == 0 && info) info[b] = inf;
}

// Small blocks (n <= 20: the 10- and 20-DOF cell blocks of XDG and p <= 2 scalar operators): G lanes
// per block (G = 16 or 32, so 64 / G blocks per wavefront), lane r of a group holds row r of its
// block in registers for the whole factorisation.  Row exchanges are bookkeeping, as in the
// workgroup-tile form: each lane carries the position its row has been given.  The pivot is the
// unused row with the largest |a| (ties to the smaller position, a NaN in position j never beaten,
// any other NaN never chosen: IDAMAX in getf2's row order), found by a shuffle argmax inside the
// group; the pivot row goes to the group's LDS slot and every lane reads it back (a group is part
// of one wavefront, whose LDS operations complete in order: no barrier).  LU: getf2's operations
// in getf2's order (multiplier a * (1 / pivot), rank-one update of the columns after j) on the
// rows after the pivot in position order.  Inverse: Gauss-Jordan in place, written out with both
// permutations applied; a zero pivot makes the group factor its block from the original values
// instead (LU and INFO).  The step loop is unrolled, so column j is a register index known at
// compile time.  No lane returns early: every shuffle has all lanes of its group taking part.
template <typename Real, int G, int NM, bool GJ>
__device__ __forceinline__ int rows_eliminate(Real (&a)[NM], int r, int n, int& pos, Real* row_slot, int* pr, int* pp, int base) {
  int inf = 0;
  bool dead = false;   // Gauss-Jordan stops at a zero pivot
#pragma unroll
  for (int j = 0; j < NM; j++) {
    if (j < n && !dead) {   // uniform over the group
      Real v = Real(-1);
      if (r < n && pos >= j) {
        v = fabs(a[j]);
        if (pos == j && v != v) v = Real(INFINITY);   // idamax: a NaN first element is never beaten
        if (!(v >= Real(0))) v = Real(-1);              // any other NaN is never a candidate
      }
      int key = v >= Real(0) ? (pos << 6) | r : 0x7fffffff;   // position first (unique), then the lane
#pragma unroll
      for (int o = G / 2; o >= 1; o >>= 1) {
        const Real ov = __shfl_xor(v, o, 64);
        const int ok = __shfl_xor(key, o, 64);
        if (ov > v || (ov == v && ok < key)) { v = ov; key = ok; }
      }
      // the row in position j is always a candidate (|a| >= 0, or a NaN made infinite), so key is set
      const int p = key & 63, ps = key >> 6;
      const Real apj = __shfl(a[j], base + p, 64);
      if (r == 0) { pr[j] = p; pp[j] = ps; }
      if (!(apj != Real(0))) {   // zero pivot (NaN passes, as in DGETF2)
        if (!inf) inf = j + 1;
        if (GJ) dead = true;
        continue;                // no exchange: the row in position j is done all the same
      }
      if (r == p) pos = j;
      else if (pos == j) pos = ps;
      const Real pinv = Real(1) / apj;
      if (r == p) {
#pragma unroll
        for (int c = 0; c < NM; c++) row_slot[c] = GJ ? (c == j ? pinv : a[c] * pinv) : a[c];
      }
      __builtin_amdgcn_wave_barrier();
      if (GJ) {
        const Real f = a[j];
        a[j] = Real(0);          // becomes -f / pivot through the update
#pragma unroll
        for (int c = 0; c < NM; c++) a[c] -= f * row_slot[c];
        if (r == p) {
#pragma unroll
          for (int c = 0; c < NM; c++) a[c] = row_slot[c];
        }
      } else if (pos > j) {      // rows after the pivot in position order
        const Real l = a[j] * pinv;
        a[j] = l;
#pragma unroll
        for (int c = j + 1; c < NM; c++) a[c] -= l * row_slot[c];
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  return inf;
}

template <typename Real, int G, int NM>
__global__ void __launch_bounds__(256) k_block_lu_rows(int nblk, int nmax, const int32_t* __restrict__ nsz, const int64_t* __restrict__ off,
                                                       Real* __restrict__ A, int32_t* __restrict__ ipiv, const int64_t* __restrict__ pivoff,
                                                       int32_t* __restrict__ info, int want_inverse) {
  __shared__ Real s_row[256 / G][NM];
  __shared__ int s_pr[256 / G][NM], s_pp[256 / G][NM];   // per step: pivot row (physical), its position before the exchange
  const int r = threadIdx.x % G, grp = threadIdx.x / G;   // row of the block this lane holds, group in the workgroup
  const int b = blockIdx.x * (256 / G) + grp;
  const int base = (threadIdx.x & 63) - r;             // first lane of the group in the wavefront
  const bool inb = b < nblk;
  const int n0 = inb ? nsz[b] : 0;
  if (inb && (n0 <= 0 || n0 > nmax) && r == 0 && info) info[b] = -1;   // size outside the launch: report
  const bool ok = inb && n0 > 0 && n0 <= nmax;
  const int n = ok ? n0 : 0;
  Real* g = ok ? A + off[b] : A;
  Real a[NM];
#pragma unroll
  for (int c = 0; c < NM; c++) a[c] = (r < n && c < n) ? g[(size_t)r * n + c] : Real(0);
  int pos = r, inf = 0;
  bool inverse = false;
  if (want_inverse) {
    inf = rows_eliminate<Real, G, NM, true>(a, r, n, pos, s_row[grp], s_pr[grp], s_pp[grp], base);
    inverse = inf == 0;
    if (!inverse) {   // singular (uniform over the group): factor the original block, LU and INFO
#pragma unroll
      for (int c = 0; c < NM; c++) a[c] = (r < n && c < n) ? g[(size_t)r * n + c] : Real(0);
      pos = r;
    }
  }
  if (!inverse) inf = rows_eliminate<Real, G, NM, false>(a, r, n, pos, s_row[grp], s_pr[grp], s_pp[grp], base);
  __builtin_amdgcn_wave_barrier();
  if (ok && r < n) {
    Real* out = g + (size_t)pos * n;
    if (inverse) {
#pragma unroll
      for (int c = 0; c < NM; c++)
        if (c < n) out[s_pr[grp][c]] = a[c];   // X[pos[r]][row placed in position c] = a[r][c]
    } else {
#pragma unroll
      for (int c = 0; c < NM; c++)
        if (c < n) out[c] = a[c];
    }
    if (ipiv) ipiv[pivoff[b] + r] = s_pp[grp][r] + 1;
  }
  if (ok && r == 0 && info) info[b] = inf;
}

// Blocks above LU_MAXN (the multi-cell blocks of Schwarz smoothers, up to LU_BIG_MAXN) factor in
// place in global memory, blocked (see the blocked forms below); the pivot search is getf2's:
// the first row with the largest |a| in the column (workgroup argmax, ties to the smaller row).
constexpr int LU_BIG_MAXN = 4096;
// the two-launch-per-panel form takes over for blocks of at least LU_SPLIT_MINN when the batch
// has fewer than LU_SPLIT_MAXBLK blocks (fewer blocks than CUs leave the one-workgroup form idle)
constexpr int LU_SPLIT_MINN = 512;
constexpr int LU_SPLIT_MAXBLK = 256;
// Returns the pivot row and, in act, whether the pivot is non-zero.  Thread 0 reads the pivot
// entry before the final barrier, so every thread decides on the same value even though the
// row swap that follows overwrites it (NaN counts as non-zero, as in DGETF2).
template <typename Real, int NT = 256>
__device__ __forceinline__ int lu_pivot_wg(const Real* a, int n, int j, Real* sv, int* si, int* s_piv, int* s_act, bool& act) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  Real best = Real(-1);
  int bi = n;
  for (int i = j + t; i < n; i += NT) {
    Real v = fabs(a[(size_t)i * n + j]);
    if (i == j && v != v) v = Real(INFINITY);   // idamax: a NaN first element is never beaten
    if (v > best || (v == best && i < bi)) { best = v; bi = i; }
  }
  wave_argmax(best, bi);
  if (lane == 0) { sv[w] = best; si[w] = bi; }
  __syncthreads();
  if (t == 0) {
    Real bv = sv[0];
    int bx = si[0];
    for (int k = 1; k < NT / 64; k++)
      if (sv[k] > bv || (sv[k] == bv && si[k] < bx)) { bv = sv[k]; bx = si[k]; }
    *s_piv = bx < n ? bx : j;   // no comparable candidate (all NaN): the diagonal, as idamax
    *s_act = a[(size_t)*s_piv * n + j] != Real(0);
  }
  __syncthreads();
  act = *s_act != 0;
  return *s_piv;
}

template <typename Real, int NT = 256>
__device__ __forceinline__ void lu_swap_rows(Real* a, int n, int r1, int r2) {
  for (int c = threadIdx.x; c < n; c += NT) {
    const Real x = a[(size_t)r1 * n + c];
    a[(size_t)r1 * n + c] = a[(size_t)r2 * n + c];
    a[(size_t)r2 * n + c] = x;
  }
}

// Blocked forms for blocks above 128 DOF (right-looking, panel width LU_NB): getf2 on the
// LU_NB columns of a panel with whole-row exchanges -- getf2's pivots, exchanges and scaling, the
// rank-1 updates applied to the panel columns only --, then U12 = L11^-1 A12 by forward
// substitution (each element gets getf2's updates in getf2's order), then the trailing update
// A22 -= L21 U12 as one rank-LU_NB product on v_mfma_f64_16x16x4f64.  Pivots, INFO and the
// factors of the panel rows are getf2's; the trailing elements receive the same products summed
// in the matrix units' order (rounding).  A column whose pivot is zero is skipped by getf2
// (no scaling, no update): its L entries count as zero in the solve and the product (zmask).
constexpr int LU_NB = 32;
typedef double lu_f64x4 __attribute__((ext_vector_type(4)));

// Panel [k0, k0 + nb) of block a (n x n, row-major), one workgroup.  Returns the panel's zero-pivot
// mask (bit jj: column k0 + jj) and sets inf to the first zero pivot (1-based) if it is still 0.
template <typename Real, int NT = 256>
__device__ __forceinline__ unsigned lu_panel(Real* a, int n, int k0, int nb, int32_t* ipiv_b, int& inf, Real* sv, int* si, int* s_piv, int* s_act) {
  const int t = threadIdx.x;
  const int kend = k0 + nb;
  unsigned zmask = 0;
  for (int j = k0; j < kend; j++) {
    bool act;
    const int p = lu_pivot_wg<Real, NT>(a, n, j, sv, si, s_piv, s_act, act);
    if (t == 0 && ipiv_b) ipiv_b[j] = p + 1;
    if (!act) {
      if (!inf) inf = j + 1;
      zmask |= 1u << (j - k0);
      __syncthreads();
      continue;
    }
    if (p != j) lu_swap_rows<Real, NT>(a, n, j, p);
    __syncthreads();
    const Real inv = Real(1) / a[(size_t)j * n + j];
    for (int i = j + 1 + t; i < n; i += NT) a[(size_t)i * n + j] *= inv;
    __syncthreads();
    const int w = kend - j - 1;   // panel columns right of j
    if (w > 0)
      for (int e = t; e < (n - j - 1) * w; e += NT) {
        const int i = j + 1 + e / w, c = j + 1 + e % w;
        a[(size_t)i * n + c] -= a[(size_t)i * n + j] * a[(size_t)j * n + c];
      }
    __syncthreads();
  }
  return zmask;
}

// U12 = L11^-1 A12 for the columns c = k0 + LU_NB + c_begin, + c_step, ... (one column per thread;
// only full panels have columns to their right).  sL: LU_NB x LU_NB staging of L11 (zero-pivot
// columns as zero), filled by the calling workgroup.
template <typename Real>
__device__ __forceinline__ void lu_trsm(Real* a, int n, int k0, unsigned zmask, int c_begin, int c_step, Real* sL) {
  for (int e = threadIdx.x; e < LU_NB * LU_NB; e += blockDim.x) {
    const int i = e / LU_NB, q = e % LU_NB;
    sL[e] = (q < i && !((zmask >> q) & 1u)) ? a[(size_t)(k0 + i) * n + k0 + q] : Real(0);
  }
  __syncthreads();
#pragma unroll 1
  for (int c = k0 + LU_NB + c_begin; c < n; c += c_step) {
    Real* pc = a + (size_t)k0 * n + c;
    Real u[LU_NB];
#pragma unroll
    for (int i = 0; i < LU_NB; i++) u[i] = pc[i * n];
    // column-oriented: u[i] -= L[i][q] u[q] for q = 0, 1, ... -- per element the order of getf2's steps
#pragma unroll
    for (int q = 0; q < LU_NB - 1; q++) {
      asm volatile("" ::: "memory");   // L11 is read column by column in each pass, not hoisted out of the loop over c
      const Real uq = u[q];
#pragma unroll
      for (int i = q + 1; i < LU_NB; i++) u[i] -= sL[i * LU_NB + q] * uq;
    }
#pragma unroll
    for (int i = 1; i < LU_NB; i++) pc[i * n] = u[i];
  }
}

// A22 -= L21 U12 on 16 x 16 tiles of the trailing matrix (rows and columns >= k0 + LU_NB, full
// panels only): each wavefront takes the tile rows ti = r_begin, + r_step, ... and keeps that
// row's L21 operands in registers while it walks the tile columns.
template <typename Real>
__device__ __forceinline__ void lu_trail(Real* a, int n, int k0, unsigned zmask, int r_begin, int r_step) {
  const int k1 = k0 + LU_NB, m = n - k1;
  if (m <= 0) return;
  const int T = (m + 15) / 16;
  const int lane = threadIdx.x & 63, col = lane & 15, grp = lane >> 4;
  const Real* U = a + (size_t)k0 * n;            // U12 rows k0 .. k0 + LU_NB - 1
#pragma unroll 1
  for (int ti = r_begin; ti < T; ti += r_step) {
    const int r0 = k1 + 16 * ti;
    Real av[LU_NB / 4];
#pragma unroll
    for (int kk = 0; kk < LU_NB / 4; kk++) {
      const int q = 4 * kk + grp;                           // panel column this lane supplies
      const int row = r0 + col;                             // A row this lane supplies
      av[kk] = (row < n && !((zmask >> q) & 1u)) ? -a[(size_t)row * n + k0 + q] : Real(0);
    }
#pragma unroll 1
    for (int tj = 0; tj < T; tj++) {
      const int c = k1 + 16 * tj + col;
      lu_f64x4 acc;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = r0 + grp + 4 * r;
        acc[r] = (row < n && c < n) ? a[(size_t)row * n + c] : Real(0);
      }
#pragma unroll
      for (int kk = 0; kk < LU_NB / 4; kk++) {
        const Real bv = c < n ? U[(4 * kk + grp) * n + c] : Real(0);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[kk], bv, acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = r0 + grp + 4 * r;
        if (row < n && c < n) a[(size_t)row * n + c] = acc[r];
      }
    }
  }
}

// One workgroup per block, every panel in turn.
template <typename Real>
__global__ void __launch_bounds__(256) k_block_lu_blocked(int nblk, int nmax, const int32_t* __restrict__ nsz, const int64_t* __restrict__ off,
                                                          Real* __restrict__ A, int32_t* __restrict__ ipiv, const int64_t* __restrict__ pivoff,
                                                          int32_t* __restrict__ info) {
  __shared__ Real sv[4];
  __shared__ int si[4];
  __shared__ int s_piv, s_act;
  __shared__ Real sL[LU_NB * LU_NB];
  const int b = blockIdx.x;
  if (b >= nblk) return;
  const int n = nsz[b];
  if (n <= 0 || n > nmax) {
    if (threadIdx.x == 0 && info) info[b] = -1;
    return;
  }
  Real* a = A + off[b];
  int32_t* ipiv_b = ipiv ? ipiv + pivoff[b] : nullptr;
  int inf = 0;
  for (int k0 = 0; k0 < n; k0 += LU_NB) {
    const int nb = n - k0 < LU_NB ? n - k0 : LU_NB;
    const unsigned zmask = lu_panel(a, n, k0, nb, ipiv_b, inf, sv, si, &s_piv, &s_act);
    if (k0 + nb < n) {   // a full panel with columns to its right
      lu_trsm(a, n, k0, zmask, (int)threadIdx.x, 256, sL);
      __syncthreads();
      lu_trail(a, n, k0, zmask, (int)(threadIdx.x >> 6), 4);
      __syncthreads();
    }
  }
  if (threadIdx.x == 0 && info) info[b] = inf;
}

// The same for few large blocks, two launches per panel.  k_lu_panel (one workgroup per block)
// factors the panel and solves U12; st[b] holds the block's INFO so far (-1: invalid size),
// st[nblk + b] the panel's zero-pivot mask.  k_lu_trail spreads the trailing product of every
// block over gridDim.x workgroups per block.
constexpr int LU_PANEL_THREADS = 512;   // few blocks: the panel's column steps get 8 wavefronts
template <typename Real>
__global__ void __launch_bounds__(LU_PANEL_THREADS) k_lu_panel(int k0, int nblk, int nmax, const int32_t* __restrict__ nsz,
                                                               const int64_t* __restrict__ off, Real* __restrict__ A,
                                                               int32_t* __restrict__ ipiv, const int64_t* __restrict__ pivoff,
                                                               int32_t* __restrict__ st) {
  __shared__ Real sv[LU_PANEL_THREADS / 64];
  __shared__ int si[LU_PANEL_THREADS / 64];
  __shared__ int s_piv, s_act;
  __shared__ Real sL[LU_NB * LU_NB];
  const int b = blockIdx.x;
  if (b >= nblk) return;
  const int n = nsz[b];
  if (k0 == 0 && threadIdx.x == 0) st[b] = (n <= 0 || n > nmax) ? -1 : 0;
  if (n <= 0 || n > nmax || k0 >= n) return;
  Real* a = A + off[b];
  int inf = k0 == 0 ? 0 : st[b];
  const int nb = n - k0 < LU_NB ? n - k0 : LU_NB;
  const unsigned zmask = lu_panel<Real, LU_PANEL_THREADS>(a, n, k0, nb, ipiv ? ipiv + pivoff[b] : nullptr, inf, sv, si, &s_piv, &s_act);
  if (k0 + nb < n) lu_trsm(a, n, k0, zmask, (int)threadIdx.x, LU_PANEL_THREADS, sL);
  if (threadIdx.x == 0) { st[b] = inf; st[nblk + b] = (int32_t)zmask; }
}

template <typename Real>
__global__ void __launch_bounds__(256) k_lu_trail(int k0, int nblk, int nmax, const int32_t* __restrict__ nsz, const int64_t* __restrict__ off,
                                                  Real* __restrict__ A, const int32_t* __restrict__ st) {
  const int b = blockIdx.y;
  if (b >= nblk) return;
  const int n = nsz[b];
  if (n <= 0 || n > nmax || k0 + LU_NB >= n) return;
  lu_trail(A + off[b], n, k0, (unsigned)st[nblk + b], (int)(blockIdx.x * 4 + (threadIdx.x >> 6)), (int)(gridDim.x * 4));
}

__global__ void k_lu_info(int nblk, const int32_t* __restrict__ st, int32_t* __restrict__ info) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nblk) info[b] = st[b];
}


// x_b <- LU_b \ x_b   (getrs with the 1-based pivots), one wavefront per block, any size up to
// 4096.  Lane l owns the rows i = l (mod 64) of x_b and is the only one to write them: the
// substitutions run column by column, the finished x_j reaches the other lanes by a shuffle from
// its owner, and each owner subtracts L_ij x_j (U_ij x_j) from its rows.  The forward sweep adds
// the terms of a row in getrs' order (k = 0 .. i-1), the backward sweep from the last column.
template <typename Real>
__global__ void __launch_bounds__(64) k_block_solve(int nblk, const int32_t* __restrict__ nsz, const int64_t* __restrict__ off,
                                                    const Real* __restrict__ LU, const int32_t* __restrict__ ipiv,
                                                    const int64_t* __restrict__ pivoff, const int64_t* __restrict__ xoff,
                                                    Real* __restrict__ x) {
  const int b = blockIdx.x;
  if (b >= nblk) return;
  const int n = nsz[b];
  if (n <= 0) return;
  const int lane = threadIdx.x;
  const Real* L = LU + off[b];
  const int32_t* pv = ipiv + pivoff[b];
  Real* xb = x + xoff[b];
  if (lane == 0)
    for (int j = 0; j < n; j++) {
      const int p = pv[j] - 1;
      if (p != j) { const Real t = xb[j]; xb[j] = xb[p]; xb[p] = t; }
    }
  __syncthreads();   // the interchanged entries are read by their owners below
  for (int j = 0; j < n; j++) {
    Real xj = lane == (j & 63) ? xb[j] : Real(0);
    xj = __shfl(xj, j & 63, 64);
    int i = lane;
    if (i <= j) i += ((j - lane) / 64 + 1) * 64;   // first owned row below j
    for (; i < n; i += 64) xb[i] -= L[(size_t)i * n + j] * xj;
  }
  for (int j = n - 1; j >= 0; j--) {
    Real xj = Real(0);
    if (lane == (j & 63)) {
      xj = xb[j] / L[(size_t)j * n + j];
      xb[j] = xj;
    }
    xj = __shfl(xj, j & 63, 64);
    for (int i = lane; i < j; i += 64) xb[i] -= L[(size_t)i * n + j] * xj;
  }
}

// y_b <- alpha * Ainv_b x_b + beta * y_b   (block-diagonal apply with explicit inverses, one wave per block)
template <typename Real>
__global__ void __launch_bounds__(64) k_block_gemv(int nblk, const int32_t* __restrict__ nsz, const int64_t* __restrict__ off,
                                                   const Real* __restrict__ Ainv, const int64_t* __restrict__ xoff, Real alpha,
                                                   const Real* __restrict__ xin, Real beta, Real* __restrict__ y) {
  const int b = blockIdx.x;
  if (b >= nblk) return;
  const int n = nsz[b];
  const Real* M = Ainv + off[b];
  const Real* xb = xin + xoff[b];
  Real* yb = y + xoff[b];
  for (int i = threadIdx.x; i < n; i += 64) {
    Real a = 0;
    for (int k = 0; k < n; k++) a += M[i * n + k] * xb[k];
    yb[i] = alpha * a + (beta != Real(0) ? beta * yb[i] : Real(0));
  }
}

}  // namespace bosss

using namespace bosss;

namespace bosss {
int ctx_device_stream(bosss_ctx* ctx, int* device, hipStream_t* stream);
}

namespace {
thread_local std::string g_lu_err;

// With a context, launch on its device (the stream argument must belong to that device);
// without one, on the calling thread's current device.
int select_device(bosss_ctx* ctx) {
  if (!ctx) return BOSSS_OK;
  int dev = 0;
  hipStream_t st = nullptr;
  if (bosss::ctx_device_stream(ctx, &dev, &st) != BOSSS_OK) { g_lu_err = "bad context"; return BOSSS_ERR_ARG; }
  if (hipSetDevice(dev) != hipSuccess) { g_lu_err = "hipSetDevice failed"; return BOSSS_ERR_HIP; }
  return BOSSS_OK;
}
}

// ---- scratch of the allocation-free factorisation (bosss_hip_block_lu_noalloc_dev) ----
// The split form keeps two status words per block between its launches.  The no-alloc entry
// point takes them from this file-scope array instead of a per-call allocation: every device the
// library runs on holds its own copy (8 KiB) until the process ends, and every call on one device
// shares it, so calls that overlap on one device race.  Every word a call reads it has written
// itself earlier in the same call (k_lu_panel writes a block's words before k_lu_trail and the
// next k_lu_panel read them), so nothing needs clearing between calls.
constexpr int LU_NOALLOC_MAXBLK = 1024;
static __device__ __align__(256) int32_t scratch[2 * LU_NOALLOC_MAXBLK];

namespace {
constexpr int LU_NOALLOC_MAXDEV = 64;
std::mutex g_scratch_mu;
int32_t* g_scratch_ptr[LU_NOALLOC_MAXDEV] = {};   // device address of each device's copy of scratch

// The address of the current device's copy of scratch, looked up once per device (a host-side
// name of a __device__ array is not a device address; hipGetSymbolAddress gives that device's).
int scratch_on_current_device(int32_t** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) { g_lu_err = std::string("hipGetDevice: ") + hipGetErrorString(e); return BOSSS_ERR_HIP; }
  if (dev < 0 || dev >= LU_NOALLOC_MAXDEV) { g_lu_err = "device index out of range for the scratch table"; return BOSSS_ERR_ARG; }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      g_lu_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return BOSSS_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<int32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return BOSSS_OK;
}
}  // namespace

// Argument checks shared by the factorisation entry points (selects ctx's device).
static int lu_check_args(bosss_ctx* ctx, int32_t nmax, const int32_t* n, const int64_t* off, const double* A, const int32_t* ipiv,
                         const int64_t* pivoff) {
  if (int rc = select_device(ctx)) return rc;
  if (!n || !off || !A || (ipiv && !pivoff)) { g_lu_err = "NULL argument"; return BOSSS_ERR_ARG; }
  if (nmax <= 0 || nmax > LU_BIG_MAXN) { g_lu_err = "block size out of range (1..4096)"; return BOSSS_ERR_ARG; }
  return BOSSS_OK;
}

// Blocks up to LU_MAXN, in registers: up to 20 DOF several blocks per wavefront, 21..128 one workgroup per block.
static int lu_launch_reg(int nblk, int nmax, const int32_t* n, const int64_t* off, double* A, int32_t* ipiv, const int64_t* pivoff,
                         int32_t* info, int want_inverse, hipStream_t s) {
  if (nmax <= 20) {
    // rows in registers, 64 / G blocks per wavefront (larger rows would not fit one lane's
    // registers: those blocks take the workgroup-tile kernel)
    if (nmax <= 16)
      hipLaunchKernelGGL((k_block_lu_rows<double, 16, 16>), dim3((unsigned)((nblk + 15) / 16)), dim3(256), 0, s, nblk, nmax, n, off, A,
                         ipiv, pivoff, info, want_inverse);
    else
      hipLaunchKernelGGL((k_block_lu_rows<double, 32, 20>), dim3((unsigned)((nblk + 7) / 8)), dim3(256), 0, s, nblk, nmax, n, off, A,
                         ipiv, pivoff, info, want_inverse);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { g_lu_err = hipGetErrorString(e); return BOSSS_ERR_HIP; }
    return BOSSS_OK;
  }
  // 21..128: one workgroup per block, the block in registers
  const dim3 grid((unsigned)nblk), blk(256);
#define LU_TILE(RA)                                                                                                     \
  if (want_inverse)                                                                                                     \
    hipLaunchKernelGGL((k_block_lu_tile<RA, true>), grid, blk, 0, s, nblk, nmax, n, off, A, ipiv, pivoff, info);        \
  else                                                                                                                  \
    hipLaunchKernelGGL((k_block_lu_tile<RA, false>), grid, blk, 0, s, nblk, nmax, n, off, A, ipiv, pivoff, info);
  if (nmax <= 32) {
    LU_TILE(2)
  } else if (nmax <= 64) {
    LU_TILE(4)
  } else if (nmax <= 112) {
    LU_TILE(7)
  } else {
    LU_TILE(8)
  }
#undef LU_TILE
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { g_lu_err = hipGetErrorString(e); return BOSSS_ERR_HIP; }
  return BOSSS_OK;
}

// The split form (two launches per panel) with its per-block status words st[2 nblk] in device
// memory the caller provides; queued on s, not synchronised.
static hipError_t lu_launch_split(int nblk, int nmax, const int32_t* n, const int64_t* off, double* A, int32_t* ipiv,
                                  const int64_t* pivoff, int32_t* info, int32_t* st, hipStream_t s) {
  for (int k0 = 0; k0 < nmax; k0 += LU_NB) {
    hipLaunchKernelGGL(k_lu_panel<double>, dim3(nblk), dim3(LU_PANEL_THREADS), 0, s, k0, nblk, nmax, n, off, A, ipiv, pivoff, st);
    const int m = nmax - k0 - LU_NB;   // trailing size of the largest block
    if (m > 0) hipLaunchKernelGGL(k_lu_trail<double>, dim3((unsigned)((m + 63) / 64), nblk), dim3(256), 0, s, k0, nblk, nmax, n, off, A, st);
  }
  if (info) hipLaunchKernelGGL(k_lu_info, dim3((nblk + 255) / 256), dim3(256), 0, s, nblk, st, info);
  return hipGetLastError();
}

extern "C" {

// Host-side error text of the block routines (the context-level message lives in bosss_hip.hip).
const char* bosss_hip_block_last_error(void) { return g_lu_err.c_str(); }

int bosss_hip_block_lu_dev(bosss_ctx* ctx, int nblk, int32_t nmax, const int32_t* n, const int64_t* off, double* A,
                           int32_t* ipiv, const int64_t* pivoff, int32_t* info, int32_t want_inverse, void* stream) {
  if (nblk <= 0) return BOSSS_OK;
  if (int rc = lu_check_args(ctx, nmax, n, off, A, ipiv, pivoff)) return rc;
  if (nmax <= LU_MAXN) return lu_launch_reg(nblk, nmax, n, off, A, ipiv, pivoff, info, want_inverse, (hipStream_t)stream);
  if (want_inverse) { g_lu_err = "inverses are limited to blocks of 128 DOF"; return BOSSS_ERR_UNSUPPORTED; }
  hipStream_t s = (hipStream_t)stream;
  if (nmax >= LU_SPLIT_MINN && nblk < LU_SPLIT_MAXBLK) {
    // few large blocks: two launches per panel, the trailing update over many workgroups
    int32_t* st = nullptr;
    hipError_t e = hipMalloc((void**)&st, 2 * (size_t)nblk * sizeof(int32_t));
    if (e != hipSuccess) { g_lu_err = hipGetErrorString(e); return BOSSS_ERR_HIP; }
    e = lu_launch_split(nblk, nmax, n, off, A, ipiv, pivoff, info, st, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);   // st is freed next
    (void)hipFree(st);
    if (e != hipSuccess) { g_lu_err = hipGetErrorString(e); return BOSSS_ERR_HIP; }
    return BOSSS_OK;
  }
  hipLaunchKernelGGL(k_block_lu_blocked<double>, dim3(nblk), dim3(256), 0, s, nblk, nmax, n, off, A, ipiv, pivoff, info);
  hipError_t eb = hipGetLastError();
  if (eb != hipSuccess) { g_lu_err = hipGetErrorString(eb); return BOSSS_ERR_HIP; }
  return BOSSS_OK;
}

int bosss_hip_block_lu_noalloc_dev(bosss_ctx* ctx, int nblk, int32_t nmax, const int32_t* n, const int64_t* off, double* A,
                                   int32_t* ipiv, const int64_t* pivoff, int32_t* info, int32_t want_inverse, void* stream) {
  if (nblk <= 0) return BOSSS_OK;
  if (int rc = lu_check_args(ctx, nmax, n, off, A, ipiv, pivoff)) return rc;
  if (nmax <= LU_MAXN) return lu_launch_reg(nblk, nmax, n, off, A, ipiv, pivoff, info, want_inverse, (hipStream_t)stream);
  if (want_inverse) { g_lu_err = "inverses are limited to blocks of 128 DOF"; return BOSSS_ERR_UNSUPPORTED; }
  if (nblk > LU_NOALLOC_MAXBLK) {
    g_lu_err = "bosss_hip_block_lu_noalloc_dev: " + std::to_string(nblk) + " blocks above " + std::to_string(LU_MAXN) +
               " DOF exceed its scratch (at most " + std::to_string(LU_NOALLOC_MAXBLK) + " per call)";
    return BOSSS_ERR_ARG;
  }
  int32_t* st = nullptr;
  if (int rc = scratch_on_current_device(&st)) return rc;
  hipError_t e = lu_launch_split(nblk, nmax, n, off, A, ipiv, pivoff, info, st, (hipStream_t)stream);
  if (e != hipSuccess) { g_lu_err = hipGetErrorString(e); return BOSSS_ERR_HIP; }
  return BOSSS_OK;
}

int bosss_hip_block_solve_dev(bosss_ctx* ctx, int nblk, const int32_t* n, const int64_t* off, const double* LU, const int32_t* ipiv,
                              const int64_t* pivoff, const int64_t* xoff, double* x, void* stream) {
  if (nblk <= 0) return BOSSS_OK;
  if (int rc = select_device(ctx)) return rc;
  if (!n || !off || !LU || !ipiv || !pivoff || !xoff || !x) { g_lu_err = "NULL argument"; return BOSSS_ERR_ARG; }
  hipLaunchKernelGGL(k_block_solve<double>, dim3(nblk), dim3(64), 0, (hipStream_t)stream, nblk, n, off, LU, ipiv,
                     pivoff, xoff, x);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { g_lu_err = hipGetErrorString(e); return BOSSS_ERR_HIP; }
  return BOSSS_OK;
}

int bosss_hip_block_gemv_dev(bosss_ctx* ctx, int nblk, const int32_t* n, const int64_t* off, const double* Ainv, const int64_t* xoff,
                             double alpha, const double* x, double beta, double* y, void* stream) {
  if (nblk <= 0) return BOSSS_OK;
  if (int rc = select_device(ctx)) return rc;
  if (!n || !off || !Ainv || !xoff || !x || !y) { g_lu_err = "NULL argument"; return BOSSS_ERR_ARG; }
  hipLaunchKernelGGL(k_block_gemv<double>, dim3(nblk), dim3(64), 0, (hipStream_t)stream, nblk, n, off, Ainv, xoff, alpha, x, beta, y);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { g_lu_err = hipGetErrorString(e); return BOSSS_ERR_HIP; }
  return BOSSS_OK;
}

// ---- host-pointer forms: the arrays are host memory, copied to the device and back on the
//      context's stream; the call returns with the results in the host arrays.
static int host_batch_check(bosss_ctx* ctx, int nblk, const int32_t* n, const int64_t* off, hipStream_t* s) {
  if (!ctx) { g_lu_err = "ctx is NULL (host-pointer calls copy on the context's stream)"; return BOSSS_ERR_ARG; }
  int dev = 0;
  if (bosss::ctx_device_stream(ctx, &dev, s) != BOSSS_OK) { g_lu_err = "bad context"; return BOSSS_ERR_ARG; }
  if (int rc = select_device(ctx)) return rc;
  if (!n || !off) { g_lu_err = "NULL argument"; return BOSSS_ERR_ARG; }
  for (int b = 0; b < nblk; b++)
    if (n[b] < 0 || off[b] < 0) { g_lu_err = "negative block size or offset"; return BOSSS_ERR_ARG; }
  return BOSSS_OK;
}
static int host_fail(hipError_t e, const char* what) {
  g_lu_err = std::string(what) + ": " + hipGetErrorString(e);
  return BOSSS_ERR_HIP;
}

int bosss_hip_block_lu(bosss_ctx* ctx, int nblk, int32_t nmax, const int32_t* n, const int64_t* off, double* A, int32_t* ipiv,
                       const int64_t* pivoff, int32_t* info, int32_t want_inverse) {
  if (nblk <= 0) return BOSSS_OK;
  hipStream_t s = nullptr;
  if (int rc = host_batch_check(ctx, nblk, n, off, &s)) return rc;
  if (!A || (ipiv && !pivoff)) { g_lu_err = "NULL argument"; return BOSSS_ERR_ARG; }
  const int64_t na = batch_extent(nblk, off, [&](int b) { return (int64_t)n[b] * n[b]; });
  const int64_t np = ipiv ? batch_extent(nblk, pivoff, [&](int b) { return (int64_t)n[b]; }) : 0;
  DevArrays d(s);
  const int32_t* dn = d.in(n, nblk);
  const int64_t* doff = d.in(off, nblk);
  double* dA = d.in((const double*)A, na);
  int32_t* dpiv = ipiv ? d.in((const int32_t*)ipiv, np) : nullptr;
  const int64_t* dpo = ipiv ? d.in(pivoff, nblk) : nullptr;
  int32_t* dinfo = info ? d.out<int32_t>(nblk) : nullptr;
  if (d.error() != hipSuccess) return host_fail(d.error(), "block_lu staging");
  if (int rc = bosss_hip_block_lu_dev(ctx, nblk, nmax, dn, doff, dA, dpiv, dpo, dinfo, want_inverse, s)) return rc;
  d.back(A, (const double*)dA, na);
  if (ipiv) d.back(ipiv, (const int32_t*)dpiv, np);
  if (info) d.back(info, (const int32_t*)dinfo, nblk);
  if (hipError_t e = d.finish()) return host_fail(e, "block_lu");
  return BOSSS_OK;
}

int bosss_hip_block_solve(bosss_ctx* ctx, int nblk, const int32_t* n, const int64_t* off, const double* LU, const int32_t* ipiv,
                          const int64_t* pivoff, const int64_t* xoff, double* x) {
  if (nblk <= 0) return BOSSS_OK;
  hipStream_t s = nullptr;
  if (int rc = host_batch_check(ctx, nblk, n, off, &s)) return rc;
  if (!LU || !ipiv || !pivoff || !xoff || !x) { g_lu_err = "NULL argument"; return BOSSS_ERR_ARG; }
  const int64_t na = batch_extent(nblk, off, [&](int b) { return (int64_t)n[b] * n[b]; });
  const int64_t np = batch_extent(nblk, pivoff, [&](int b) { return (int64_t)n[b]; });
  const int64_t nx = batch_extent(nblk, xoff, [&](int b) { return (int64_t)n[b]; });
  DevArrays d(s);
  const int32_t* dn = d.in(n, nblk);
  const int64_t* doff = d.in(off, nblk);
  const double* dLU = d.in(LU, na);
  const int32_t* dpiv = d.in(ipiv, np);
  const int64_t* dpo = d.in(pivoff, nblk);
  const int64_t* dxo = d.in(xoff, nblk);
  double* dx = d.in((const double*)x, nx);
  if (d.error() != hipSuccess) return host_fail(d.error(), "block_solve staging");
  if (int rc = bosss_hip_block_solve_dev(ctx, nblk, dn, doff, dLU, dpiv, dpo, dxo, dx, s)) return rc;
  d.back(x, (const double*)dx, nx);
  if (hipError_t e = d.finish()) return host_fail(e, "block_solve");
  return BOSSS_OK;
}

int bosss_hip_block_gemv(bosss_ctx* ctx, int nblk, const int32_t* n, const int64_t* off, const double* Ainv, const int64_t* xoff,
                         double alpha, const double* x, double beta, double* y) {
  if (nblk <= 0) return BOSSS_OK;
  hipStream_t s = nullptr;
  if (int rc = host_batch_check(ctx, nblk, n, off, &s)) return rc;
  if (!Ainv || !xoff || !x || !y) { g_lu_err = "NULL argument"; return BOSSS_ERR_ARG; }
  const int64_t na = batch_extent(nblk, off, [&](int b) { return (int64_t)n[b] * n[b]; });
  const int64_t nx = batch_extent(nblk, xoff, [&](int b) { return (int64_t)n[b]; });
  DevArrays d(s);
  const int32_t* dn = d.in(n, nblk);
  const int64_t* doff = d.in(off, nblk);
  const double* dAi = d.in(Ainv, na);
  const int64_t* dxo = d.in(xoff, nblk);
  const double* dx = d.in(x, nx);
  double* dy = d.in((const double*)y, nx);
  if (d.error() != hipSuccess) return host_fail(d.error(), "block_gemv staging");
  if (int rc = bosss_hip_block_gemv_dev(ctx, nblk, dn, doff, dAi, dxo, alpha, dx, beta, dy, s)) return rc;
  d.back(y, (const double*)dy, nx);
  if (hipError_t e = d.finish()) return host_fail(e, "block_gemv");
  return BOSSS_OK;
}

}  // extern "C"
