// Strip-pipelined exact Gauss-Seidel sweeps (device code), shared by the
// momentum block's sequential preconditioners (momentum.hip) and the Poisson
// block's symmetric Gauss-Seidel (poisson.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace covof {
namespace {

// Exact lexicographic Gauss-Seidel sweeps of a 5-point block on the GPU: the
// SSOR block solves of apply_preconditioner (src/module_linear_solvers.f90:
// 137-316, ssor_solve_u / _v / _p :1278-1403) and the SP Poisson block's
// symmetric Gauss-Seidel (src/module_simulation.f90:2709-2758).
//
// Lexicographic Gauss-Seidel is order dependent: cell (i,j)
// needs the new W and S values and the old E and N ones.  Here one lane owns
// one row of a 64-row strip and walks it along i, one cell per step, lane l
// l steps behind lane l-1: the new W value is the lane's own last result, the
// new S value the lower lane's last result (a lane shuffle; across strips an
// LDS slot or, between passes, global memory), and the old E / N / own values
// come from the half-sweep's input buffer, so every cell sees exactly the
// operands of the lexicographic order and the result is the sequential
// recurrence value for value.  A workgroup of W waves runs W strips at once,
// wave w SEQ_LAG steps behind wave w-1 (one barrier per batch of steps);
// strips beyond W follow in further passes.  The backward half mirrors it
// (strips from the top, lane 0 on the strip's top row).
//
// The block's arrays are kept in a skewed order -- strip kk, step s = i-1+l,
// lane l -- so that the 64 lanes of a step read 64 consecutive doubles: all
// loads of a step are coalesced, and they are issued SEQ_D steps ahead.
// arithmetic of a cell: the SSOR block solves, the Poisson SGS, the SVOF
// Helmholtz smoothing (smooth_svof_helmholtz, src/module_interface_vof.f90:
// 387-476)
enum { SEQ_SSOR = 0, SEQ_POISSON_SGS = 1, SEQ_SVOF = 2 };
constexpr int SEQ_D = 4;
constexpr int SEQ_MAXW = 8;
// waves synchronise once per batch of 2 SEQ_D steps; wave w runs SEQ_LAG
// steps behind wave w-1, so the value a strip's first row needs from the
// strip below was made SEQ_LAG - 63 = 17 steps before: before the barrier
// that precedes even the earliest load of it (loads run up to 3 SEQ_D - 1
// steps ahead of their step), so it is read from a ring of SEQ_RING LDS
// slots per wave together with the step's other operands
constexpr int SEQ_LAG = 64 + 4 * SEQ_D;
constexpr int SEQ_RING = 32;

struct SeqBlk {
  int ni, nj, K, S;             // block size, strips, steps per strip (ni + 63)
  double *rec;                  // per cell (skewed): a_P, a_W, a_E, a_S, a_N, rhs (8-double records)
  double *xa, *xb;              // state buffers (skewed): input / output of a half-sweep
  double *hW, *hE, *hS, *hN;    // halo snapshots x(0,j), x(ni+1,j), x(i,0), x(i,nj+1)
};
// record fields
enum { SR_P = 0, SR_W = 1, SR_E = 2, SR_S = 3, SR_N = 4, SR_R = 5 };

// A workgroup barrier that orders LDS only: the per-step barrier of the
// sweeps must not wait for the global loads issued ahead (a full
// __syncthreads() drains them); the strips exchange values through LDS only.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// lane i <- lane i-1 of a double (DPP wave_shr:1; lane 0 keeps its own value)
__device__ __forceinline__ double wave_shr1(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false),
                          __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xf, 0xf, false));
}

__host__ __device__ __forceinline__ long sk_at(int S, int kk, int s, int l) { return ((long)kk * S + s) * 64 + l; }
__host__ __device__ __forceinline__ long sk_of(int S, int i, int j) {
  const int kk = (j - 1) >> 6, l = (j - 1) & 63;
  return sk_at(S, kk, i - 1 + l, l);
}


// The operands of one cell, loaded ahead: coefficients, rhs, own old value,
// the two old neighbours, and the new neighbour across rows where it comes
// from a halo or an earlier pass.
struct SeqOps {
  double P, W, E, S, N, R, XO, O1, O2, B2, RG;
};

// One half-sweep of the strips (FWD: forward) from xi into xo.  A lane's
// row is walked as t = 0 .. ni-1 (FWD: i = t+1; backward: i = ni-t), lane l
// l steps after lane 0, wave w SEQ_LAG w steps after wave 0; everything that
// stays fixed along the row is set once per pass.
template <bool FWD, int MODE>
__device__ void seq_half(const SeqBlk &b, const double *__restrict__ xi, double *__restrict__ xo, double omega,
                         int W, double (*ring)[SEQ_RING]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int l = FWD ? lane : 63 - lane;  // row of the strip
  const int ni = b.ni, nj = b.nj, S = b.S;
  const long SS = (long)S * 64;  // one strip
  const int passes = (b.K + W - 1) / W;
  const int nsteps = ((S + SEQ_LAG * (W - 1)) + 2 * SEQ_D - 1) / (2 * SEQ_D) * (2 * SEQ_D);
  const long dstep = FWD ? 64 : -64;  // storage offset of the next cell along the row
  const long pad = (long)b.K * SS + lane;
  // across rows: offset of the old neighbour (FWD: N, backward: S) within the
  // block, and of the new one in the neighbouring strip (lane 0 only)
  const long offX = FWD ? (l < 63 ? 65 : SS - 63 * 64 - 63) : (l > 0 ? -65 : -SS + 63 * 64 + 63);
  const long offY = FWD ? -SS + 63 * 64 + 63 : SS - 63 * 64 - 63;
  const double *hX = FWD ? b.hN : b.hS, *hY = FWD ? b.hS : b.hN;  // halo rows, by i - 1
  for (int pass = 0; pass < passes; ++pass) {
    const int kq = pass * W + w;
    const int kk = FWD ? kq : b.K - 1 - kq;
    const int j = 64 * kk + l + 1;
    const bool row_on = w < W && kq < b.K && j <= nj;
    const long at0 = row_on ? sk_at(S, kk, FWD ? l : ni - 1 + l, l) : 0;  // storage of t = 0
    const bool cross_in = FWD ? j < nj : j > 1;          // old neighbour across rows inside the block
    const bool edge_row = FWD ? j == 1 : j == nj;        // new neighbour across rows from the halo
    const bool prev_pass = lane == 0 && w == 0 && !edge_row;  // ... or from the strip of an earlier pass
    const double bA = row_on ? (FWD ? b.hW[j - 1] : b.hE[j - 1]) : 0.0;  // new neighbour along the row at t = 0
    const double eA = row_on ? (FWD ? b.hE[j - 1] : b.hW[j - 1]) : 0.0;  // old neighbour along the row at t = ni-1
    // branch-free: every step issues the same loads (an inactive lane reads
    // index 0), so they stay in flight SEQ_D steps
    auto fetch = [&](int g, SeqOps &q) {
      const int t = g - SEQ_LAG * w - lane;
      const bool act = row_on && (unsigned)t < (unsigned)ni;
      const long at = act ? at0 + t * dstep : 0;
      const int i = FWD ? t + 1 : ni - t;
      const double2 *r = (const double2 *)(b.rec + at * 8);
      const double2 r0 = r[0], r1 = r[1], r2 = r[2];
      q.P = r0.x, q.W = r0.y, q.E = r1.x, q.S = r1.y, q.N = r2.x, q.R = r2.y;
      q.XO = xi[at];
      if (MODE != SEQ_POISSON_SGS) {
        q.O1 = xi[(act && t < ni - 1) ? at + dstep : 0];
        q.O2 = *(act ? (cross_in ? xi + at + offX : hX + (i - 1)) : xi);
      } else {
        q.O1 = q.O2 = 0.0;
      }
      q.B2 = *(act ? (edge_row ? hY + (i - 1) : (prev_pass ? xo + at + offY : xi)) : xi);
      q.RG = ring[w > 0 ? w - 1 : 0][(g - (SEQ_LAG - 63)) & (SEQ_RING - 1)];  // the strip below's row
    };
    double cur = 0.0;
    auto step = [&](int g, const SeqOps &o) {
      const int t = g - SEQ_LAG * w - lane;
      const bool act = row_on && (unsigned)t < (unsigned)ni;
      const double up = wave_shr1(cur);     // the previous row's last result
      const double nA = t == 0 ? bA : cur;  // new neighbour along the row
      double nB;                            // new neighbour across rows
      if (edge_row || prev_pass) nB = o.B2;
      else if (lane > 0) nB = up;
      else nB = o.RG;
      double nv;
      if (MODE == SEQ_SSOR) {
        const double oA = t == ni - 1 ? eA : o.O1;  // old neighbour along the row
        const double xW = FWD ? nA : oA, xE = FWD ? oA : nA;
        const double xS = FWD ? nB : o.O2, xN = FWD ? o.O2 : nB;
        double dg = o.P;
        if (fabs(dg) < 1.0e-30) dg = 1.0;
        double val = o.R;
        if (o.W != 0.0) val = val + o.W * xW;
        if (o.E != 0.0) val = val + o.E * xE;
        if (o.S != 0.0) val = val + o.S * xS;
        if (o.N != 0.0) val = val + o.N * xN;
        nv = (1.0 - omega) * o.XO + omega * (val / dg);
      } else if (MODE == SEQ_SVOF) {  // (C + sum of the neighbours) / a_P, terms E, W, N, S
        const double oA = t == ni - 1 ? eA : o.O1;
        const double xW = FWD ? nA : oA, xE = FWD ? oA : nA;
        const double xS = FWD ? nB : o.O2, xN = FWD ? o.O2 : nB;
        const double sum = o.E * xE + o.W * xW + o.N * xN + o.S * xS;
        nv = (o.R + sum) / o.P;
      } else if (FWD) {  // Poisson SGS forward: (D + L) z = r, the wrap terms where present
        double val = o.R;
        if (o.W != 0.0) val = val + o.W * nA;
        if (o.S != 0.0) val = val + o.S * nB;
        nv = val / o.P;
      } else {  // backward: (D + U) z = D z_forward
        double val = o.P * o.XO;
        if (o.E != 0.0) val = val + o.E * nA;
        if (o.N != 0.0) val = val + o.N * nB;
        nv = val / o.P;
      }
      // an inactive lane stores into the padding slot past the strips
      xo[act ? at0 + t * dstep : pad] = nv;
      cur = act ? nv : cur;
      if (W > 1 && lane == 63 && act) ring[w][g & (SEQ_RING - 1)] = nv;
    };
    // two batches of SEQ_D steps in flight: batch k+1 loads while batch k computes
    SeqOps qa[SEQ_D], qb[SEQ_D];
#pragma unroll
    for (int d = 0; d < SEQ_D; ++d) fetch(d, qa[d]);
    for (int g0 = 0; g0 < nsteps; g0 += 2 * SEQ_D) {
#pragma unroll
      for (int d = 0; d < SEQ_D; ++d) fetch(g0 + SEQ_D + d, qb[d]);
#pragma unroll
      for (int d = 0; d < SEQ_D; ++d) step(g0 + d, qa[d]);
#pragma unroll
      for (int d = 0; d < SEQ_D; ++d) fetch(g0 + 2 * SEQ_D + d, qa[d]);
#pragma unroll
      for (int d = 0; d < SEQ_D; ++d) step(g0 + SEQ_D + d, qb[d]);
      if (W > 1) lds_barrier();
    }
    __syncthreads();  // the strips of this pass are read by the next one
  }
}

// halo snapshot of the state x (skewed) as the single-rank exchange sets it:
// x(0,j) = x(ni,j), x(ni+1,j) = x(1,j) with periodic x, x(i,0) = x(i,nj),
// x(i,nj+1) = x(i,1) with periodic y; zero on other sides
// (neumann: a physical side copies the adjacent interior value instead)
__device__ void seq_snapshot(bool perx, bool pery, const SeqBlk &b, const double *x, bool neumann = false) {
  for (int t = threadIdx.x; t < b.nj; t += blockDim.x) {
    b.hW[t] = perx ? x[sk_of(b.S, b.ni, t + 1)] : (neumann ? x[sk_of(b.S, 1, t + 1)] : 0.0);
    b.hE[t] = perx ? x[sk_of(b.S, 1, t + 1)] : (neumann ? x[sk_of(b.S, b.ni, t + 1)] : 0.0);
  }
  for (int t = threadIdx.x; t < b.ni; t += blockDim.x) {
    b.hS[t] = pery ? x[sk_of(b.S, t + 1, b.nj)] : (neumann ? x[sk_of(b.S, t + 1, 1)] : 0.0);
    b.hN[t] = pery ? x[sk_of(b.S, t + 1, 1)] : (neumann ? x[sk_of(b.S, t + 1, b.nj)] : 0.0);
  }
  __syncthreads();
}

// n_sweeps symmetric sweeps from zero (SSOR: relaxation omega, ssor_solve_*;
// the Poisson mode: one symmetric Gauss-Seidel sweep): with a periodic side
// the halo is exchanged before every sweep but the first and between the
// halves.  The SVOF mode starts from the caller's b.xa and fills the halo
// (zero-gradient copies, then the exchange) before every sweep only.  The
// result ends in b.xa.  One workgroup of up to SEQ_MAXW waves.
template <int MODE>
__device__ void seq_sweeps(bool perx, bool pery, const SeqBlk &b, int n_sweeps, double omega) {
  __shared__ double ring[SEQ_MAXW][SEQ_RING];
  const bool sync = perx || pery;
  const int W = min((int)(blockDim.x >> 6), b.K);
  const long nsk = (long)b.K * b.S * 64;
  if (MODE != SEQ_SVOF) {
    for (long k = threadIdx.x; k < nsk; k += blockDim.x) b.xa[k] = 0.0;
    for (int t = threadIdx.x; t < b.nj; t += blockDim.x) b.hW[t] = b.hE[t] = 0.0;
    for (int t = threadIdx.x; t < b.ni; t += blockDim.x) b.hS[t] = b.hN[t] = 0.0;
  }
  __syncthreads();
  for (int sweep = 1; sweep <= n_sweeps; ++sweep) {
    if (MODE == SEQ_SVOF) seq_snapshot(perx, pery, b, b.xa, true);
    else if (sync && sweep > 1) seq_snapshot(perx, pery, b, b.xa);
    seq_half<true, MODE>(b, b.xa, b.xb, omega, W, ring);
    if (MODE != SEQ_SVOF && sync) seq_snapshot(perx, pery, b, b.xb);
    seq_half<false, MODE>(b, b.xb, b.xa, omega, W, ring);
  }
}

// the size of a block's skewed arrays (+ the padding slot of inactive lanes)
inline long seq_len(int ni, int nj) { return (long)((nj + 63) / 64) * (ni + 63) * 64 + 64; }

// one wave per strip of the largest of the blocks, at most SEQ_MAXW
inline int seq_waves(int nj) {
  const int K = (nj + 63) / 64;
  return K < SEQ_MAXW ? K : SEQ_MAXW;
}

}  // namespace
}  // namespace covof
