// Batched sparse Newton-Raphson on gfx950: one wavefront per instance runs the whole Newton loop.
//
// Per Newton step, for each instance (reference path: SparseNewtonRaphsonPowerFlow.RunNewton, :213-262):
//   1. injections, mismatch and Jacobian (FP64, FDLIBM sincos once per Y cell): ComputeInjections :269-289, residual
//      :218-233, RefillJacobian :342-380. First every Y cell's two injection terms and its off-diagonal Jacobian
//      entries, one cell per lane; then per bus the ordered sums, the diagonal entries and the mismatch;
//   2. the verdicts: converged / diverged / iteration cap (:235-238, NewtonShell.DivergedBeyondLimit :177-178);
//   3. Gilbert-Peierls refactor on the plan (SparseLuFactorization.Refactor :246-306), level-parallel: the columns of
//      one dependency level are independent, so up to 8 of them are factored side by side (8 lanes each), each in its
//      own LDS workspace, with static pivots and a growth check (a failed check pauses the instance for a fresh CPU
//      symbolic pass);
//   4. triangular solves (Solve :338-371) in the LU precision, level-scheduled over rows with the vector in LDS, each
//      row pulling its updates in the order the reference's column sweeps apply them; then FP64 iterative refinement
//      r = f - J dx on the FP64 Jacobian;
//   5. the damped update (NewtonShell.ApplyStep :128-136).
// Compiled with -ffp-contract=off: in FP64 every value is the reference's arithmetic, operation for operation.
//
// Every phase is a chain of dependent memory reads (pattern, then the value it points at), so the loops keep the
// reads for later iterations in flight in registers (software pipelining) instead of waiting on each in turn. All
// global pointers are cast to the global address space: generic (flat) loads would also count against the LDS wait
// counter, and every LDS wait would then drain the reads in flight.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "fdlibm.h"
#include "gridsim_hip.h"
#include "nr_kernels.h"

#define GLOBAL __attribute__((address_space(1)))
#define LDSP __attribute__((address_space(3)))

namespace gs {

constexpr int kWave = 64;

// native vector types (HIP's vector classes do not carry an address space)
typedef int v2i __attribute__((ext_vector_type(2)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

template <typename T>
__device__ __forceinline__ GLOBAL T* glob(T* p) { return (GLOBAL T*)p; }
template <typename V>
__device__ __forceinline__ GLOBAL const V* globv(const void* p) { return (GLOBAL const V*)p; }
template <typename T>
__device__ __forceinline__ LDSP T* lds(T* p) { return (LDSP T*)p; }

__device__ __forceinline__ double nan_max(double a, double b) {
  if (a != a || b != b) return __builtin_nan("");
  return a > b ? a : b;
}

__device__ __forceinline__ double wave_max(double v) {
  for (int off = kWave / 2; off > 0; off >>= 1) v = nan_max(v, __shfl_xor(v, off, kWave));
  return v;
}

// A lane's value from another lane of its 16-lane row by a DPP pattern (every lane has a source; no LDS traffic):
// quad_perm [1,0,3,2] and [2,3,0,1] swap neighbours and pairs within each quad, row_half_mirror (lane 7 - i of each
// 8-lane half) and row_mirror (lane 15 - i of each row) reach the other half.
constexpr int kDppQuadXor1 = 0xB1, kDppQuadXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false); }
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)dpp_i32<CTRL>((int)b), hi = (unsigned)dpp_i32<CTRL>((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// maximum over the tl lanes of one track (aligned groups of tl, a power of two the same in every lane). Tracks of 8 and
// 16 lanes (nearly every pivot row) reduce with DPP moves: after the two quad steps each quad holds its maximum, and
// the mirrors pair each lane with one of the other quad (half) of its group. Wider tracks shuffle through LDS. The
// maximum of non-negative values and NaN does not depend on the order it is taken in.
__device__ __forceinline__ double track_max(double v, int tl) {
  if (tl <= 16) {
    v = nan_max(v, dpp_f64<kDppQuadXor1>(v));
    v = nan_max(v, dpp_f64<kDppQuadXor2>(v));
    v = nan_max(v, dpp_f64<kDppHalfMirror>(v));
    if (tl == 16) v = nan_max(v, dpp_f64<kDppMirror>(v));
    return v;
  }
  for (int off = tl / 2; off > 0; off >>= 1) v = nan_max(v, __shfl_xor(v, off, kWave));
  return v;
}

__device__ __forceinline__ int track_any(int v, int tl) {
  if (tl <= 16) {
    v |= dpp_i32<kDppQuadXor1>(v);
    v |= dpp_i32<kDppQuadXor2>(v);
    v |= dpp_i32<kDppHalfMirror>(v);
    if (tl == 16) v |= dpp_i32<kDppMirror>(v);
    return v;
  }
  for (int off = tl / 2; off > 0; off >>= 1) v |= __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ int wave_any(int v) { return __any(v) ? 1 : 0; }

// The lane index, opaque to the optimiser: passed to each phase of the Newton loop so that lane-dependent addresses
// are computed inside the phase rather than hoisted out of the loop (where they would hold registers throughout).
__device__ __forceinline__ int opaque_lane(int lane) {
  asm volatile("" : "+v"(lane));
  return lane;
}

// Ordering of LDS traffic between the lanes of one wavefront (the hardware executes a wave's LDS instructions in
// order; this stops the compiler from moving them).
#define WAVE_SYNC() __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront")
// Global-memory writes of this wavefront made visible to its own later reads (work-group scope: same CU, shared L1).
#define WG_SYNC() __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup")

// Waits for this wave's LDS (and scalar-memory) operations only. Placed at the end of a rarely taken branch whose LDS
// reads may otherwise still be pending where the branches join: the compiler would then wait there on every path,
// including the frequent one, which never issued them. (s_waitcnt lgkmcnt(0), vmcnt and expcnt left at their maxima.)
__device__ __forceinline__ void lds_drained() { __builtin_amdgcn_s_waitcnt(0xC07F); }

// One instance's buffers, as global-memory pointers.
template <typename T>
struct Inst {
  GLOBAL const double *pspec, *qspec, *vm_start, *va_start;
  GLOBAL double *vm, *va, *p, *q, *f, *dx, *r, *jv;
  GLOBAL v2d* tpq;
  GLOBAL T *vals, *fac, *dval, *xg;   // fac: the factor part of vals (vals + n_a)
  GLOBAL const Patch* patch;
  GLOBAL const v2d* ygb;   // the Y values (g, b) per cell: the network's, or the instance's own
  int n_patch;
  int pidx[kMaxPatch];   // patched Y cells (-1: unused)
};

// Y value of cell k: the base value, or this instance's patch of it.
template <typename T>
__device__ __forceinline__ void y_value(const DevPlan& P, const Inst<T>& I, int k, double g0, double b0, double* g,
                                        double* b) {
  *g = g0;
  *b = b0;
#pragma unroll
  for (int t = 0; t < kMaxPatch; t++)
    if (I.pidx[t] == k) { *g = I.patch[t].g; *b = I.patch[t].b; }
}

typedef float v4f __attribute__((ext_vector_type(4)));

// The four Jacobian values of one Y cell (or one bus's diagonal block) in the LU precision, to four consecutive
// positions of vals: one contiguous store per lane. The per-instance values each Newton step writes in whole lines
// (Jacobian cells, injection terms, factors: consecutive lanes' stores adjacent) are stored non-temporal (NT):
// hundreds of instances per XCD far exceed its L2, so keeping them there would only evict the plan's record streams,
// which every instance reads.
template <typename T, bool NT = true>
__device__ __forceinline__ void store4(GLOBAL T* p, double h, double nn, double mm, double l) {
  if constexpr (sizeof(T) == 8) {
    GLOBAL v2d* q = (GLOBAL v2d*)p;
    if constexpr (NT) {
      __builtin_nontemporal_store((v2d){h, nn}, q);
      __builtin_nontemporal_store((v2d){mm, l}, q + 1);
    } else {
      q[0] = (v2d){h, nn};
      q[1] = (v2d){mm, l};
    }
  } else {
    const v4f v = {(float)h, (float)nn, (float)mm, (float)l};
    if constexpr (NT) __builtin_nontemporal_store(v, (GLOBAL v4f*)p);
    else *(GLOBAL v4f*)p = v;
  }
}

// The FP64 CSR copy of the Jacobian the refinement residual and the iterative fallback read (scattered 8-byte stores).
__device__ __forceinline__ void put_csr(GLOBAL double* jv, const v4i& s, double h, double nn, double mm, double l) {
  if (s.x >= 0) jv[s.x] = h;
  if (s.y >= 0) jv[s.y] = nn;
  if (s.z >= 0) jv[s.z] = mm;
  if (s.w >= 0) jv[s.w] = l;
}

// ------------------------------------------------------------------------------------------------ mismatch + J
// VALS: the Jacobian goes, in the LU precision, to the A part of vals in Y-cell order (cell k's lane writes H, N, M, L
// to 4k .. 4k + 3: those of cell k, or with P.y_tr those of its transpose; bus i's diagonal block goes to its own
// cell's place; the refactor's ASSIGN rows gather them from there); KEEP: also to the FP64 CSR copy (without it, the
// CSR positions are not even loaded). The lanes' A positions (aslots, -1 for an entry the Jacobian does not have) say
// which values count towards max |J|.
// Every Y cell's pattern and base value are two loads: (row, column) and (g, b). Both passes walk blocks of 64 cells (buses): blocks
// blk0, blk0 + stride, ... of the wave (one wave: 0, 1; a level-parallel grid: one block per work-group).

// Where the cells pass puts each Y cell's two injection terms and the bus pass reads them back: the instance's tpq
// array in HBM (the two passes apart), or the wave's LDS scratch for one block of whole rows (the passes interleaved
// block by block, mismatch_blocked).
struct TermsGlobal {
  GLOBAL v2d* t;
  __device__ void put(int k, double tp, double tq) const { __builtin_nontemporal_store((v2d){tp, tq}, &t[k]); }
  __device__ v2d get(int k) const { return t[k]; }
};
struct TermsLds {
  LDSP double* p;
  LDSP double* q;
  int base;   // the block's first cell
  __device__ void put(int k, double tp, double tq) const { p[k - base] = tp; q[k - base] = tq; }
  __device__ v2d get(int k) const { return (v2d){p[k - base], q[k - base]}; }
};

// The two injection terms and off-diagonal Jacobian entries of the Y cells [k_lo, k_hi), one cell per lane. Returns
// whether a Jacobian value of this lane was not finite.
template <typename T, bool KEEP, bool VALS = true, typename TS = TermsGlobal>
__device__ __forceinline__ int mismatch_cells(const DevPlan& P, const Inst<T>& I, int lane, int k_lo, int k_hi,
                                              int blk0, int stride, const TS& terms, double* jmax) {
  GLOBAL const v2i* y_ij = globv<v2i>(P.y_ij);
  GLOBAL const v2d* y_gb = I.ygb;
  GLOBAL const v4i* slots = globv<v4i>(P.slots);
  GLOBAL const v4i* aslots = globv<v4i>(P.aslots);
  const int nblk = (k_hi - k_lo + kWave - 1) / kWave;
  int lbad = 0;
  double jm = 0.0;
  // stage 0 holds the pattern of block t + 2 (lane-strided), stage 1 the gathered state of block t + 1
  const int iters = blk0 < nblk ? (nblk - blk0 + stride - 1) / stride : 0;
  int i1 = 0, j1 = 0, i2 = 0, j2 = 0;
  double g1 = 0, b1 = 0, g2 = 0, b2 = 0;
  v4i s1 = (v4i)(-1), a1 = s1, s2 = s1, a2 = s1;
  double vai1 = 0, vaj1 = 0, vmi1 = 0, vmj1 = 0;
  auto load_pattern = [&](int t, int& i, int& j, double& g, double& b, v4i& s, v4i& a) {
    const int k = k_lo + lane + (blk0 + t * stride) * kWave;
    const int kk = k < k_hi ? k : 0;
    const v2i ij = y_ij[kk];
    const v2d gb = y_gb[kk];
    i = ij.x; j = ij.y; g = gb.x; b = gb.y; a = aslots[kk];
    if (KEEP) s = slots[kk];
  };
  // (cells past the end read cell 0: every load is in bounds and unconditional, so the waits stay partial)
  load_pattern(0, i1, j1, g1, b1, s1, a1);
  load_pattern(1, i2, j2, g2, b2, s2, a2);
  vai1 = I.va[i1]; vaj1 = I.va[j1]; vmi1 = I.vm[i1]; vmj1 = I.vm[j1];
  for (int t = 0; t < iters; t++) {
    // this cell
    const int k = k_lo + lane + (blk0 + t * stride) * kWave;
    const int i = i1, j = j1;
    const double g0 = g1, b0 = b1, vai = vai1, vaj = vaj1, vmi = vmi1, vmj = vmj1;
    const v4i s = s1, a = a1;
    // advance the pipeline: gather for block t + 1, pattern for block t + 2
    i1 = i2; j1 = j2; g1 = g2; b1 = b2; s1 = s2; a1 = a2;
    vai1 = I.va[i1]; vaj1 = I.va[j1]; vmi1 = I.vm[i1]; vmj1 = I.vm[j1];
    load_pattern(t + 2, i2, j2, g2, b2, s2, a2);
    if (k < k_hi) {
      double g, b;
      y_value(P, I, k, g0, b0, &g, &b);
      double sn, cs;
      gs_fdlibm::sincos(vai - vaj, &sn, &cs);
      const double tp = vmj * (g * cs + b * sn);
      const double tq = vmj * (g * sn - b * cs);
      terms.put(k, tp, tq);
      if (j != i) {
        // cell (r, c)'s entries from V_r, V_c and the sine of va_r - va_c: this cell's (r = i), or with y_tr its
        // transpose's (r = j): the Y value is the same bit for bit, and sin(va_j - va_i) is -sin(va_i - va_j) exactly
        // (an odd sine and an exact difference), or for a zero difference that difference itself, signed as it is
        const bool tr = P.y_tr != 0;
        const double vr = tr ? vmj : vmi, vc = tr ? vmi : vmj;
        const double st = !tr ? sn : sn == 0.0 ? vaj - vai : 0.0 - sn;
        const double h = vr * vc * (g * st - b * cs);
        const double nn = vr * (g * cs + b * st);
        const double mm = -vr * vc * (g * cs + b * st);
        const double l = vr * (g * st - b * cs);
        if (KEEP) put_csr(I.jv, s, h, nn, mm, l);
        if (VALS) store4<T>(I.vals + 4 * (size_t)k, h, nn, mm, l);
        lbad |= !(isfinite(h) && isfinite(nn) && isfinite(mm) && isfinite(l));
        jm = a.x >= 0 && fabs(h) > jm ? fabs(h) : jm;
        jm = a.y >= 0 && fabs(nn) > jm ? fabs(nn) : jm;
        jm = a.z >= 0 && fabs(mm) > jm ? fabs(mm) : jm;
        jm = a.w >= 0 && fabs(l) > jm ? fabs(l) : jm;
      }
    }
  }
  *jmax = jm;
  return lbad;
}

// Per bus the ordered sums of its row's terms, the diagonal Jacobian entries and the mismatch; the next bus's
// scalars are loaded ahead; buses [i_lo, i_hi). Returns this lane's max |f| (NaN-propagating); *bad: a Jacobian value
// was not finite.
template <typename T, bool KEEP, bool VALS = true, typename TS = TermsGlobal>
__device__ __forceinline__ double mismatch_buses(const DevPlan& P, const Inst<T>& I, int lane, int i_lo, int i_hi,
                                                 int blk0, int stride, const TS& terms, int* bad, double* jmax) {
  GLOBAL const v2d* y_gb = I.ygb;
  GLOBAL const v4i* slots = globv<v4i>(P.slots);
  GLOBAL const v4i* aslots = globv<v4i>(P.aslots);
  GLOBAL const int* y_rp = glob(P.y_rp);
  GLOBAL const int* y_diag = glob(P.y_diag);
  GLOBAL const int* ppos = glob(P.ppos);
  GLOBAL const int* qpos = glob(P.qpos);
  double lmax = 0.0, jm = 0.0;
  int lbad = 0;
  int i = i_lo + lane + blk0 * kWave;
  const int step = stride * kWave;
  int nk0 = 0, nk1 = 0, nkd = 0, npp = -1, nqq = -1;
  double nvm = 0;
  auto load_bus = [&](int ii) {
    const int c = ii < i_hi ? ii : 0;
    nk0 = y_rp[c]; nk1 = y_rp[c + 1]; nkd = y_diag[c]; npp = ppos[c]; nqq = qpos[c]; nvm = I.vm[c];
  };
  load_bus(i);
  for (; i < i_hi; i += step) {
    const int k0 = nk0, k1 = nk1, kd = nkd, pp = npp, qq = nqq;
    const double vmi = nvm;
    load_bus(i + step);
    // (the diagonal cell's base value and slots repeated by bus after the cells': coalesced loads)
    const v2d gbd = y_gb[P.nnz_y + i];
    const double gd0 = gbd.x, bd0 = gbd.y;
    const v4i s = KEEP ? slots[P.nnz_y + i] : (v4i)(-1), a = aslots[P.nnz_y + i];
    const double ps = I.pspec[i], qs = I.qspec[i];
    double pi = 0.0, qi = 0.0;
    int k = k0;
    for (; k + 2 <= k1; k += 2) {
      const v2d u = terms.get(k), v = terms.get(k + 1);
      pi += u.x; qi += u.y;
      pi += v.x; qi += v.y;
    }
    if (k < k1) { const v2d u = terms.get(k); pi += u.x; qi += u.y; }
    const double P_i = vmi * pi, Q_i = vmi * qi;
    I.p[i] = P_i;
    I.q[i] = Q_i;
    double g, b;
    y_value(P, I, kd, gd0, bd0, &g, &b);
    const double h = -Q_i - b * vmi * vmi;
    const double nn = P_i / vmi + g * vmi;
    const double mm = P_i - g * vmi * vmi;
    const double l = Q_i / vmi - b * vmi;
    if (KEEP) put_csr(I.jv, s, h, nn, mm, l);
    // (plain stores: the lanes' diagonal blocks lie in their buses' stretches, one partial line each, which the L2
    // merges with the neighbouring cells' values; streamed past it, each would cost a partial write of its own)
    if (VALS) store4<T, false>(I.vals + 4 * (size_t)kd, h, nn, mm, l);
    lbad |= !(isfinite(h) && isfinite(nn) && isfinite(mm) && isfinite(l));
    jm = a.x >= 0 && fabs(h) > jm ? fabs(h) : jm;
    jm = a.y >= 0 && fabs(nn) > jm ? fabs(nn) : jm;
    jm = a.z >= 0 && fabs(mm) > jm ? fabs(mm) : jm;
    jm = a.w >= 0 && fabs(l) > jm ? fabs(l) : jm;
    if (pp >= 0) { const double fv = ps - P_i; I.f[pp] = fv; lmax = nan_max(lmax, fabs(fv)); }
    if (qq >= 0) { const double fv = qs - Q_i; I.f[qq] = fv; lmax = nan_max(lmax, fabs(fv)); }
  }
  *bad = lbad;
  *jmax = jm;
  return lmax;
}

// Both passes interleaved over the plan's blocks of whole rows (P.mb_ptr for FP64, mb_ptr4 for FP32: consecutive
// buses, at most 64, whose rows hold at most as many cells as the scratch takes; blocks blk0, blk0 + stride, ...):
// the cells pass puts the block's injection terms in the wave's LDS scratch (scr: the refactor workspace, free here),
// and the bus pass sums them from there, so the terms never leave the CU. Returns this lane's max |f|
// (NaN-propagating); *bad: a Jacobian value was not finite; *jmax: this lane's max |J|.
template <typename T, bool KEEP, bool VALS = true>
__device__ __forceinline__ double mismatch_blocked(const DevPlan& P, const Inst<T>& I, int lane, int blk0, int stride,
                                                   LDSP double* scr, int* bad, double* jmax) {
  constexpr bool kF32 = sizeof(T) == 4;
  GLOBAL const int* mb = glob(kF32 ? P.mb_ptr4 : P.mb_ptr);
  const int nmb = kF32 ? P.n_mb4 : P.n_mb;
  const int cap = ws_slots(P) * (int)sizeof(T) / 16;   // cells per block at most: two FP64 terms each
  GLOBAL const int* y_rp = glob(P.y_rp);
  int lbad = 0;
  double jm = 0.0, lmax = 0.0;
  for (int b = blk0; b < nmb; b += stride) {
    const int i0 = mb[b], i1 = mb[b + 1];
    const int c0 = y_rp[i0], c1 = y_rp[i1];
    const TermsLds ts{scr, scr + cap, c0};
    double jc, jb;
    int bb;
    lbad |= mismatch_cells<T, KEEP, VALS>(P, I, lane, c0, c1, 0, 1, ts, &jc);
    WAVE_SYNC();
    lmax = nan_max(lmax, mismatch_buses<T, KEEP, VALS>(P, I, lane, i0, i1, 0, 1, ts, &bb, &jb));
    lbad |= bb;
    jm = jc > jm ? jc : jm;
    jm = jb > jm ? jb : jm;
    WAVE_SYNC();   // (the next block rewrites the scratch: this block's reads first)
  }
  *bad = lbad;
  *jmax = jm;
  return lmax;
}

// The two passes apart, the terms through HBM (the iterative fallback's kernel). Returns max |f| (NaN-propagating) over
// the wave; sets *bad when a Jacobian value is not finite.
// *jnorm: the largest |entry| of the Jacobian (entries that are finite).
template <typename T, bool KEEP, bool VALS = true>
__device__ __forceinline__ double mismatch_jacobian(const DevPlan& P, const Inst<T>& I, int lane, int* bad,
                                                   double* jnorm) {
  const TermsGlobal tg{I.tpq};
  double jc, jb;
  const int cbad = mismatch_cells<T, KEEP, VALS>(P, I, lane, 0, P.nnz_y, 0, 1, tg, &jc);
  WG_SYNC();
  int bbad;
  const double lmax = mismatch_buses<T, KEEP, VALS>(P, I, lane, 0, P.n, 0, 1, tg, &bbad, &jb);
  *bad = wave_any(cbad | bbad);
  *jnorm = wave_max(jc > jb ? jc : jb);
  return wave_max(lmax);
}

// ------------------------------------------------------------------------------------------------ refactor
// Level-chunked left-looking Gilbert-Peierls on the plan. Track g of a chunk (lanes 8g..8g+7) owns one column and the
// workspace ws[g * wstride, ...): slots [0, nu - 1) its off-diagonal U rows in topological order, slot nu - 1 the
// pivot, slots [nu, nu + nl - 1) its L rows. Every workspace slot receives its updates in the reference's order.
//   EXACT = false (Newton): a static pivot is kept while |pivot| >= growth_tol * max |candidate| and |pivot| >
//   abs_floor (kPivotFloorRel x max |J|, above every row's own floor).
//   EXACT = true (seam): the reference's own pivot search must pick the plan's row (first strict maximum over the
//   candidates in reach order), above that row's floor; anything else refuses, as Refactor returns null.
// Returns -1 on success, else a failing pivot step.
//
// The refactor runs the plan's unified row stream: each row is one operation of the wavefront (ASSIGN a chunk's A
// entries into the workspace, UPDATE, PIVOT check, WRITE the factors out) with one record per lane. Every row's record
// is loaded kRing rows ahead and its value (an A entry or an L entry from vals) kAhead rows ahead; the plan places no-op
// rows so that no row reads an L column whose level ended fewer than kAhead rows before it, so each value is final when
// it is loaded and nothing is ever loaded twice. The ring registers are only rewritten at fixed points of the unrolled
// loop, so the compiler's wait counts stay partial and the loads stay in flight. The stream is padded to whole rings.
//
// Each WRITE row stores its 64 lanes' slots to 64 consecutive factor positions of vals (L divided by the pivot, U and
// the pivot as they are; an idle lane to a position nothing reads), one contiguous store; the later columns' updates
// and the triangular solves read the factors there (a track's 8 lanes read 8 consecutive slots of one column, one or
// two memory segments), and the pivots go to dval by pivot step as well. A chunk's first write row is done by its
// PIVOT row (kRsPivotWrite), each lane's kind taken from the pivot record and its position from the stream's count of
// factor-writing rows (P.rs_wr0 at a level-parallel chunk's first row).
constexpr int kRing = kStreamPad, kHalf = kRing / 2, kAhead = kRsValueAhead;
static_assert(kAhead < kRing, "the value of a row is loaded after its record");

// Rows [row0, row1) of the stream (the whole stream, or with CHUNK one chunk of a level-parallel run): the loop runs
// whole rings from row0's pair and, with CHUNK, treats the rows outside [row0, row1) as no-ops (the whole stream's
// rounding rows are no-ops already).
template <typename T, bool EXACT, bool CHUNK = false>
__device__ __forceinline__ int refactor(const DevPlan& P, GLOBAL T* vals, GLOBAL T* dval, LDSP T* ws,
                                        double growth_tol, double abs_floor,
                                        GLOBAL const double* rowfloor, int lane, double* min_ratio, int row0,
                                        int row1) {
  // rows go in pairs (one 16-byte record per lane for two rows): start at the even row of row0's pair, the rows
  // before row0 doing nothing
  const int r0e = row0 & ~1;
  GLOBAL const v4i* rs_pair = globv<v4i>(P.rs_rec) + (size_t)(r0e / 2) * kWave;
  GLOBAL const int* piv_pos = glob(P.piv_pos);
  GLOBAL const int* prow = glob(P.prow);
  const int wsz = kTracks * P.wstride;
  const int first = row0 - r0e, nrows = row1 - r0e;
  double mratio = *min_ratio;
  // The row's kind is read from the lane's own record (every lane holds the same), not from a per-row array: a value
  // the compiler knows to be uniform would be moved to a scalar register as soon as it is loaded, which waits for it.
  v4i rp[kRing / 2];
  T rv[kRing];
  auto rec_of = [&](int s) -> v2i {
    const v4i q = rp[s / 2];
    return (s % 2) ? (v2i){q.z, q.w} : (v2i){q.x, q.y};
  };
  // The rows without a value (no-op, pivot, write) load entry 0 in a batch run: the same line for every lane, where
  // their own field (a pivot step, a factor position) would touch 64 scattered lines of the value array, which costs
  // more than the select when many waves share a CU. A level-parallel chunk (one wave on its CU, latency-bound) loads
  // at the field unselected: every field indexes inside vals (vals_slots).
  auto value_index = [](const v2i& r) {
    if constexpr (CHUNK) return r.y & kRsIndexMask;
    const int k = (int)((unsigned)r.y >> kRsMetaShift) & kRsKindMask;
    return k == kRsAssign || k == kRsUpdate ? (r.y & kRsIndexMask) : 0;
  };
  // The first loads in the order the loop's rows issue them (each odd row reloads its pair; from row kRing - kAhead on
  // each row also loads the value of the row kAhead back), each pinned in place: the wait counts the compiler derives
  // at the loop head merge this path with the loop's back edge, so any other order here would make every pass of the
  // loop wait for nearly all loads in flight.
#pragma unroll
  for (int s = 0; s < kRing; s++) {
    if (s % 2 == 1) rp[s / 2] = rs_pair[(size_t)(s / 2) * kWave + lane];
    if (s >= kRing - kAhead) {
      const int h = s - (kRing - kAhead);
      rv[h] = vals[value_index(rec_of(h))];
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  T piv = T(1);
  int failed = -1;
  int write_row = 0;   // WRITE rows since the chunk's PIVOT row (a WRITE row's slots follow from it)
  int n_wr = CHUNK ? P.rs_wr0[row0] : 0;   // the stream's factor-writing rows before this one
  // (a failed pivot check ends the loop at the next ring boundary, not at once: an exit from inside the unrolled body
  // would make the compiler drain every load in flight at each row; the rows run meanwhile only write factors that
  // the caller discards)
  for (int base = 0; base < nrows && failed < 0; base += kRing) {
#pragma unroll
    for (int s = 0; s < kRing; s++) {
      const v2i r4 = rec_of(s);
      // Every lane holds the row's kind and flags: taken as a scalar here, at the row itself (the record arrived
      // kRing rows ago), so the dispatch below is a scalar branch and the row's lanes are never masked.
      int meta = (int)((unsigned)__builtin_amdgcn_readfirstlane(r4.y) >> kRsMetaShift);
      if (CHUNK && (base + s >= nrows || base + s < first)) meta = kRsNop;
      const int kind = meta & kRsKindMask;
      const int wsh = 3 + ((meta >> kRsWidthShift) & 3), TLr = 1 << wsh;   // lanes per track of the row's chunk
      if (kind == kRsUpdate) {
        // a lane without an update works on its own spare slot
        const int dst = r4.x & 0xFFFF, src = (int)((unsigned)r4.x >> 16);
        ws[dst] = ws[dst] - rv[s] * ws[src];
        WAVE_SYNC();
      } else if (kind == kRsAssign) {
        if (meta & kRsClear) {
          for (int q = lane; q < wsz; q += kWave) ws[q] = T(0);
          WAVE_SYNC();
        }
        ws[r4.x] = rv[s];   // (a lane without an A entry: its spare slot)
        WAVE_SYNC();
      } else if (kind == kRsPivot) {
        // pivot and growth check, one track per TLr lanes (all lanes of a track are active or idle together)
        int fail = -1;
        write_row = 0;
        if (r4.x != -1) {
          const int k = r4.y & kRsIndexMask, ps = r4.x & 0xFFFF;
          const int lo = ps + 1, hi = lo + (int)((unsigned)r4.x >> 16);
          piv = ws[ps];
          const double pabs = fabs((double)piv);
          double cand = pabs;
          int moved = 0;
          const int before = EXACT ? piv_pos[k] : 0;
          for (int c = lo + (lane & (TLr - 1)); c < hi; c += TLr) {
            const double cabs = fabs((double)ws[c]);
            cand = nan_max(cand, cabs);
            if (EXACT) moved |= (c - lo < before) ? !(cabs < pabs) : (cabs > pabs);
          }
          const double best = track_max(cand, TLr);
          bool bad;
          if (EXACT) {
            bad = track_any(moved, TLr) || !(pabs > rowfloor[prow[k]]);
          } else {
            const double ratio = pabs / best;
            bad = !(ratio >= growth_tol) || !(pabs > abs_floor) || !(pabs > 0.0) || !isfinite(pabs);
            mratio = ratio < mratio ? ratio : mratio;
          }
          if (bad) fail = k;
        }
        // any failing track fails the factorisation (the first failure's lowest lane's column reported)
        const unsigned long long fmask = __ballot(fail >= 0);
        if (fmask && failed < 0) failed = __shfl(fail, __ffsll((long long)fmask) - 1, kWave);
        if (meta & kRsPivotWrite) {
          // the chunk's write row 0: slot t = lane % TLr of the lane's track; its kind from the pivot record (U before
          // the pivot slot, the pivot, L up to the last candidate, idle past it or on an idle track), its position the
          // stream's next 64
          const int tb = (lane >> wsh) * P.wstride, t = lane & (TLr - 1);
          const int pt = r4.x == -1 ? -1 : (r4.x & 0xFFFF) - tb, nc = (int)((unsigned)r4.x >> 16);
          const int wk = pt < 0 ? 0 : t < pt ? kWriteU : t == pt ? kWritePivot : t <= pt + nc ? kWriteL : 0;
          const T v = ws[tb + t];
          const T lv = v / piv;
          __builtin_nontemporal_store(wk == kWriteL ? lv : v, &vals[P.n_a + 64 * n_wr + lane]);
          if (wk == kWritePivot) dval[r4.y & kRsIndexMask] = v;
          write_row = 1;
          n_wr++;
        }
        lds_drained();
      } else if (kind == kRsWrite) {
        // every lane reads its slot (an idle lane a slot inside the workspace), forms the L quotient and stores to
        // its own factor position: the row's 64 stores are one contiguous piece
        const int wk = (int)((unsigned)r4.x >> 24), pos = r4.y & kRsIndexMask;
        const T v = ws[(lane >> wsh) * P.wstride + write_row * TLr + (lane & (TLr - 1))];
        const T lv = v / piv;
        __builtin_nontemporal_store(wk == kWriteL ? lv : v, &vals[pos]);
        if (wk == kWritePivot) dval[r4.x & kRsIndexMask] = v;
        write_row++;
        n_wr++;
        lds_drained();
      }
      // the level's L columns are written: publish them before any later row loads them
      if (meta & kRsLevelEnd) WG_SYNC();
      // refill this slot with row base + s + kRing; load the value of row base + s + kHalf
      // after the second row of a pair: the pair kRing rows on
      if (s % 2 == 1) rp[s / 2] = rs_pair[(size_t)((base + s) / 2 + kRing / 2) * kWave + lane];
      // the value of row base + s + kAhead (the rows without one load an entry that is always there)
      const int hs = (s + kAhead) % kRing;
      rv[hs] = vals[value_index(rec_of(hs))];
    }
  }
  WG_SYNC();
  mratio = -wave_max(-mratio);
  *min_ratio = mratio;
  return failed;
}

// ------------------------------------------------------------------------------------------------ solves
// out = A^{-1} rhs with the factors. Rows of one level are independent: a pass takes up to 64 rows of one level, one
// per lane. Its PRODUCT steps form the products L_ij x_j (U_ij x_j) of the pass's entries, 64 per step, into the scratch
// area (each lane's row's entries contiguous, in the order the reference's column sweeps apply them: forward columns
// ascending, back columns descending); its CHAIN step then has each lane subtract its row's products from the row's
// value one by one in that order (back: then divide by the pivot). A product rounds the same whoever forms it, so
// every entry is the reference's value, and a row costs one CHAIN iteration per entry (LDS reads and one subtraction)
// instead of one step. Passes larger than the scratch area run in segments; a lane's running value stays in its
// register across them. The step records are loaded kRing steps ahead and the factor values kHalf steps ahead (the
// factors are final before a solve starts; the streams are padded by kRing steps); a PRODUCT step gathers its values
// from where the refactor stored them (a WRITE row's stores are contiguous, the solves' loads scattered: a store
// costs the vector-memory path more than a load). A solve vector in LDS is
// read at the step itself; one in global memory is loaded kHalf steps ahead with the values (x[row] for a CHAIN step,
// x[col] for a PRODUCT step), except at PRODUCT steps the plan marks kProductLate, whose columns end too shortly
// before them. XP: the solve vector (LDS or global); scr: the scratch area (LDS, the wave's refactor workspace).
// Steps [step0, step1) of the stream (the whole stream, or one pass of a level-parallel run): whole rings from step0,
// the steps from step1 on doing nothing.
// Maximum of a non-negative int over the wave, with DPP moves (no LDS traffic): shifts within rows of 16 lanes, then
// the row broadcasts carry each row's maximum up; lane 63 ends with the maximum. Lanes a move has no source for read 0.
__device__ __forceinline__ int wave_max_nonneg(int v) {
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true));   // row_shr:1
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, true));   // row_shr:2
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, true));   // row_shr:4
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, true));   // row_shr:8
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, true));   // row_bcast:15
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, true));   // row_bcast:31
  return __builtin_amdgcn_readlane(v, 63);
}

template <typename T, bool BACK, bool XLDS, typename XP, bool WIDE = false>
__device__ __forceinline__ void solve_stream(GLOBAL const v2i* rec, int step0, int step1, GLOBAL const T* vals,
                                             GLOBAL const T* dval, XP x, LDSP T* scr, int lane) {
  rec += (size_t)step0 * kWave;
  const int steps = step1 - step0;
  v2i rr[kRing];
  T rv[kRing], rd[kRing];
  // the kind of a step (every lane holds it), and the value index / pivot row a lane's record names for it
  // a PRODUCT step's value index, a CHAIN step's row (its pivot); in a batch run the other steps load entry 0, one
  // line for every lane, rather than 64 scattered ones; a level-parallel pass loads values at the field unselected
  // (every field indexes inside the factors: vals_slots) and the pivot at a CHAIN step's row only
  auto kind_of = [](const v2i& r) { return (int)((unsigned)r.y >> kRsMetaShift) & kStepKindMask; };
  auto value_index = [&](const v2i& r) {
    if constexpr (WIDE) return r.y & kRsIndexMask;
    return kind_of(r) == kStepProduct ? (r.y & kRsIndexMask) : 0;
  };
  auto pivot_index = [&](const v2i& r) { return kind_of(r) == kStepChain ? (r.y & kRsIndexMask) : 0; };
  // the entry of x a step reads: a CHAIN step's row, a PRODUCT step's column (an idle lane's -1, and the padding's,
  // read as 0)
  auto x_index = [&](const v2i& r) { return kind_of(r) == kStepChain ? (r.y & kRsIndexMask) : max(r.x, 0); };
  static_assert(kHalf == kStepXAhead, "x loads ahead by the distance the plan marks late steps for");
  constexpr bool XPRE = !XLDS;
  T rx[kRing];
  // first loads in the loop's own order (see refactor): records of steps 0..kHalf-1, then per step the record of step
  // kHalf+h and the values of step h; the loop's first kHalf steps load the values of steps kHalf..kRing-1
#pragma unroll
  for (int s = 0; s < kRing; s++) {
    rr[s] = rec[(size_t)s * kWave + lane];
    if (s >= kRing - kHalf) {
      const int h = s - (kRing - kHalf);
      rv[h] = vals[value_index(rr[h])];
      rd[h] = BACK ? dval[pivot_index(rr[h])] : T(1);
      if constexpr (XPRE) rx[h] = x[x_index(rr[h])];
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  T acc = T(0);
  int sbase = 0;   // scratch slot of this PRODUCT step's lane 0
  for (int base = 0; base < steps; base += kRing) {
#pragma unroll
    for (int s = 0; s < kRing; s++) {
      const int t = base + s;
      const v2i r2 = rr[s];
      // the step's kind is the same in every lane: a scalar, taken at the step itself
      const int meta = (int)((unsigned)__builtin_amdgcn_readfirstlane(r2.y) >> kRsMetaShift);
      int kind = meta & kStepKindMask;
      if (t >= steps) kind = 0;
      if (kind == kStepProduct) {
        const int c = r2.x;
        T xv;
        if (XPRE && !(meta & kProductLate)) {
          xv = rx[s];
        } else {
          xv = x[c >= 0 ? c : 0];
          // the load completes inside this branch, so that where the branches join the other loads in flight are
          // waited for only as far as the prefetched values need (vmcnt 0, expcnt and lgkmcnt left at their maxima)
          if constexpr (XPRE) __builtin_amdgcn_s_waitcnt(0x0F70);
        }
        scr[sbase + lane] = rv[s] * xv;   // (an idle lane's slot is never read)
        sbase += kWave;
        WAVE_SYNC();
      } else if (kind == kStepChain) {
        const int row = r2.y & kRsIndexMask, fl = (int)((unsigned)r2.y >> kRsMetaShift);
        const bool active = !(fl & kChainIdle);
        const int st = r2.x & 0xFFFF, len = (int)((unsigned)r2.x >> 16);
        const int maxlen = wave_max_nonneg(len);
        const T x0 = XPRE ? rx[s] : x[row];
        if (fl & kChainBegin) acc = x0;
#pragma unroll 4
        for (int e = 0; e < maxlen; e++) {
          const T pv = scr[st + (e < len ? e : 0)];
          acc = e < len ? acc - pv : acc;
        }
        if ((fl & kChainEnd) && active) x[row] = BACK ? acc / rd[s] : acc;
        if (XLDS) WAVE_SYNC(); else WG_SYNC();
        sbase = 0;
      }
      rr[s] = rec[(size_t)(t + kRing) * kWave + lane];
      // the factor values of step t + kHalf: a PRODUCT step's entries, a back CHAIN step's pivots (the other steps
      // load an entry that is always there)
      const int hs = (s + kHalf) % kRing;
      rv[hs] = vals[value_index(rr[hs])];
      if (BACK) rd[hs] = dval[pivot_index(rr[hs])];
      if constexpr (XPRE) rx[hs] = x[x_index(rr[hs])];
    }
  }
}

// Loops over a per-instance array with a dependent gather per entry take kBatch entries per thread at a time, every
// load issued before any is used: one entry at a time each would pay the full memory latency. Thread t of nt takes
// entries t, t + nt, ...

// x[pinv[r]] = rhs[r] (the solve vector from the right-hand side, row-permuted)
template <typename XP, typename T>
__device__ __forceinline__ void permute_in(XP x, GLOBAL const int* pinv, GLOBAL const double* rhs, int m, int t,
                                           int nt) {
  constexpr int kBatch = 8;
  for (int r0 = t; r0 < m; r0 += kBatch * nt) {
    int pr[kBatch];
    double rv[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      const int r = r0 + u * nt < m ? r0 + u * nt : 0;
      pr[u] = pinv[r];
      rv[u] = rhs[r];
    }
#pragma unroll
    for (int u = 0; u < kBatch; u++)
      if (r0 + u * nt < m) x[pr[u]] = (T)rv[u];
  }
}

// This thread's max |dx[k]| and whether one is not finite
__device__ __forceinline__ double dx_max(GLOBAL const double* dx, int m, int t, int nt, int* bad) {
  constexpr int kBatch = 8;
  double lmax = 0.0;
  int lbad = 0;
  for (int k0 = t; k0 < m; k0 += kBatch * nt) {
    double v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; u++) v[u] = dx[k0 + u * nt < m ? k0 + u * nt : 0];
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      if (k0 + u * nt >= m) continue;
      lbad |= !isfinite(v[u]);
      lmax = fabs(v[u]) > lmax ? fabs(v[u]) : lmax;
    }
  }
  *bad = lbad;
  return lmax;
}

// ApplyStep: va += damp dx(P), vm += damp dx(Q) floored at 0.01, each bus's unknowns at ppos / qpos of dx
template <typename T>
__device__ __forceinline__ void apply_step(const Inst<T>& I, GLOBAL const int* ppos, GLOBAL const int* qpos,
                                           double damp, int n, int t, int nt) {
  constexpr int kBatch = 8;
  for (int i0 = t; i0 < n; i0 += kBatch * nt) {
    int pp[kBatch], qq[kBatch];
    double va[kBatch], vm[kBatch], dp[kBatch], dq[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      const int i = i0 + u * nt < n ? i0 + u * nt : 0;
      pp[u] = ppos[i];
      qq[u] = qpos[i];
      va[u] = I.va[i];
      vm[u] = I.vm[i];
    }
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      dp[u] = I.dx[pp[u] >= 0 ? pp[u] : 0];
      dq[u] = I.dx[qq[u] >= 0 ? qq[u] : 0];
    }
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      const int i = i0 + u * nt;
      if (i >= n) continue;
      if (pp[u] >= 0) I.va[i] = va[u] + damp * dp[u];
      if (qq[u] >= 0) {
        const double v = vm[u] + damp * dq[u];
        I.vm[i] = v < 0.01 ? 0.01 : v;
      }
    }
  }
}

// OUTK: out in pivot-step order (out[k] = x[k], one contiguous store per 64 entries) instead of the unknowns' own order
// (out[q[k]] = x[k]).
template <typename T, bool XLDS, typename XP, bool OUTK = false>
__device__ __forceinline__ void lu_solve(const DevPlan& P, GLOBAL const T* fac, GLOBAL const T* dval,
                                         GLOBAL const double* rhs, GLOBAL double* out, XP x,
                                         LDSP T* scr, int lane) {
  const int m = P.m;
  GLOBAL const int* pinv = glob(P.pinv);
  GLOBAL const int* q = glob(P.q);
  permute_in<XP, T>(x, pinv, rhs, m, lane, kWave);
  if (XLDS) WAVE_SYNC(); else WG_SYNC();
  solve_stream<T, false, XLDS>(globv<v2i>(P.fs_rec), 0, P.fs_steps, fac, dval, x, scr, lane);
  solve_stream<T, true, XLDS>(globv<v2i>(P.bs_rec), 0, P.bs_steps, fac, dval, x, scr, lane);
  for (int k = lane; k < m; k += kWave) out[OUTK ? k : q[k]] = (double)x[k];
  WG_SYNC();
}

// r = f - J dx (FP64, CSR rows; each row's products summed in column order); KORDER: dx in pivot-step order.
// Rows t, t + nt, ... (a wave: its lane and 64; a work-group: its thread and size).
template <bool KORDER = false>
__device__ __forceinline__ void residual(const DevPlan& P, GLOBAL const double* jv, GLOBAL const double* f,
                                         GLOBAL const double* dx, GLOBAL double* r, int t, int nt = kWave) {
  GLOBAL const int* j_rp = glob(P.j_rp);
  GLOBAL const int* j_ci = glob(KORDER ? P.j_ci_k : P.j_ci);
  const int m = P.m;
  int i = t;
  int nk0 = 0, nk1 = 0;
  double nf = 0;
  auto load_row = [&](int ii) {
    const int c = ii < m ? ii : 0;
    nk0 = j_rp[c]; nk1 = j_rp[c + 1]; nf = f[c];
  };
  load_row(i);
  for (; i < m; i += nt) {
    const int k0 = nk0, k1 = nk1;
    const double fi = nf;
    load_row(i + nt);
    double s = 0.0;
    int k = k0;
    for (; k + 4 <= k1; k += 4) {
      const int c0 = j_ci[k], c1 = j_ci[k + 1], c2 = j_ci[k + 2], c3 = j_ci[k + 3];
      const double v0 = jv[k], v1 = jv[k + 1], v2 = jv[k + 2], v3 = jv[k + 3];
      const double x0 = dx[c0], x1 = dx[c1], x2 = dx[c2], x3 = dx[c3];
      s += v0 * x0; s += v1 * x1; s += v2 * x2; s += v3 * x3;
    }
    for (; k < k1; k++) s += jv[k] * dx[j_ci[k]];
    r[i] = fi - s;
  }
  WG_SYNC();
}

// ------------------------------------------------------------------------------------------------ the kernel
// The pair's solve vector: one wave at a time. Lane 0 takes the LDS lock word with a compare-and-swap and the wave
// spins (sleeping between tries) until it has it, at most kPairSpinLimit tries; the holder never waits on the other
// wave, so the lock always comes free. Taking it orders the holder's LDS accesses after the previous holder's.
__device__ __forceinline__ bool take_lock(LDSP int* lock, int lane) {
  for (int tries = 0; tries < kPairSpinLimit; tries++) {
    int got = 0;
    if (lane == 0) {
      int expected = 0;
      got = __hip_atomic_compare_exchange_strong(lock, &expected, 1, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    if (__shfl(got, 0, kWave)) return true;
    __builtin_amdgcn_s_sleep(4);
  }
  return false;
}

__device__ __forceinline__ void give_lock(LDSP int* lock, int lane) {
  WAVE_SYNC();
  if (lane == 0) __hip_atomic_store(lock, 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// At least two waves per SIMD (registers capped at 256): 8 instances resident per CU when LDS allows.
template <typename T, bool XLDS, bool PAIR>
__global__ __launch_bounds__(kWave * (PAIR ? kPairWaves : 1), 2) void nr_batch_kernel(
    const InstDesc* __restrict__ descs, const DevPlan P, const NrParams prm) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x % kWave;
  const int wave = PAIR ? (int)(threadIdx.x / kWave) : 0;
  const int inst = PAIR ? (int)blockIdx.x * kPairWaves + wave : (int)blockIdx.x;
  // LDS: each wave's refactor workspace, then (when it fits) the triangular-solve vector, shared by the pair, then
  // the pair's lock word
  LDSP T* ws = lds(reinterpret_cast<T*>(smem)) + wave * ws_slots(P);
  LDSP T* xl = lds(reinterpret_cast<T*>(smem)) + (PAIR ? kPairWaves : 1) * ws_slots(P);
  LDSP int* x_lock = reinterpret_cast<LDSP int*>(xl + P.m);
  if constexpr (PAIR) {
    if (threadIdx.x == 0) *x_lock = 0;
    __syncthreads();   // the only work-group barrier: from here on the two waves run independently
    if (inst >= prm.n_inst) return;
  }
  const InstDesc& d = descs[inst];
  Inst<T> I;
  I.pspec = glob(d.pspec); I.qspec = glob(d.qspec); I.vm_start = glob(d.vm_start); I.va_start = glob(d.va_start);
  I.vm = glob(d.vm); I.va = glob(d.va); I.p = glob(d.p); I.q = glob(d.q);
  I.f = glob(d.f); I.dx = glob(d.dx); I.r = glob(d.r); I.jv = glob(d.jv);
  I.tpq = (GLOBAL v2d*)d.tpq;
  I.vals = glob(reinterpret_cast<T*>(d.vals));
  I.fac = I.vals + P.n_a;
  I.dval = glob(reinterpret_cast<T*>(d.dval));
  I.xg = glob(reinterpret_cast<T*>(d.xg));
  I.patch = glob(&descs[inst].patch[0]);
  I.ygb = d.y_gb_own ? globv<v2d>(d.y_gb_own) : globv<v2d>(P.y_gb);
  I.n_patch = d.n_patch;
#pragma unroll
  for (int t = 0; t < kMaxPatch; t++) I.pidx[t] = t < I.n_patch ? d.patch[t].idx : -1;
  InstState* st = d.state;
  const int n = P.n, m = P.m;

  // a fresh instance (the launch starts every instance, or this one was added to a running batch) starts from its
  // start point with its counters reset; any other finished or paused one returns at once
  const int was = prm.init ? 0 : st->done;
  const bool fresh = prm.init || was == kDoneFresh;
  if (fresh || prm.load_start) {
    for (int i0 = lane; i0 < n; i0 += 4 * kWave) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int i = i0 + u * kWave < n ? i0 + u * kWave : 0;
        a[u] = I.vm_start[i];
        b[u] = I.va_start[i];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (i0 + u * kWave < n) { I.vm[i0 + u * kWave] = a[u]; I.va[i0 + u * kWave] = b[u]; }
    }
    WG_SYNC();
  }
  if (!fresh && was) return;
  int iter = fresh ? 0 : st->iter;
  double start_mis = fresh ? __builtin_nan("") : st->start_mis;
  double min_ratio = fresh ? 1.0 : st->min_ratio;
  double last_from = fresh ? __builtin_nan("") : st->last_from;
  int step_failures = fresh ? 0 : st->step_failures;
  int converged = 0, diverged = 0, paused = 0, fail_col = -1, steps = 0;
  double max_mis = __builtin_nan("");
  uint64_t t_mis = fresh ? 0 : st->t_mismatch, t_fac = fresh ? 0 : st->t_refactor;
  uint64_t t_sol = fresh ? 0 : st->t_solve, t_upd = fresh ? 0 : st->t_update;
  uint64_t clk = wall_clock64();
  auto tick = [&](uint64_t& acc) {
    const uint64_t now = wall_clock64();
    acc += now - clk;
    clk = now;
  };
  int lock_timeout = 0;
  auto solve = [&](GLOBAL const double* rhs, GLOBAL double* out) {
    if constexpr (PAIR) {
      if (!take_lock(x_lock, lane)) { lock_timeout = 1; return; }
      lu_solve<T, true, LDSP T*, true>(P, I.fac, I.dval, rhs, out, xl, ws, opaque_lane(lane));
      give_lock(x_lock, lane);
    } else if constexpr (XLDS) {
      lu_solve<T, true, LDSP T*, true>(P, I.fac, I.dval, rhs, out, xl, ws, opaque_lane(lane));
    } else {
      lu_solve<T, false, GLOBAL T*, true>(P, I.fac, I.dval, rhs, out, I.xg, ws, opaque_lane(lane));
    }
  };
  for (;;) {
    // the launch's steps are taken: pause before the next mismatch (the next launch starts with it)
    if (steps >= prm.steps_this_launch) { paused = 1; break; }
    int bad;
    double jnorm;
    {
      // blocked (the plan's blocks fit every row; the workspace is free between refactors)
      LDSP double* mscr = reinterpret_cast<LDSP double*>(ws);
      double jm;
      max_mis = wave_max(prm.refine_steps > 0
                             ? mismatch_blocked<T, true>(P, I, opaque_lane(lane), 0, 1, mscr, &bad, &jm)
                             : mismatch_blocked<T, false>(P, I, opaque_lane(lane), 0, 1, mscr, &bad, &jm));
      bad = wave_any(bad);
      jnorm = wave_max(jm);
    }
    WG_SYNC();
    tick(t_mis);
    if (iter == 0) start_mis = max_mis;
    if (max_mis < prm.tolerance) { converged = 1; break; }
    if (max_mis > nan_max(prm.diverge_limit, start_mis)) { diverged = 1; break; }
    if (iter >= prm.max_iterations) break;
    if (bad || !isfinite(max_mis)) break;  // NewtonShell.SolveStep refuses a non-finite J or f
    // (a plan the host made at this very state is the factorisation's own: its first step is not checked)
    const bool trust = prm.trust_first && steps == 0;
    fail_col = refactor<T, false>(P, I.vals, I.dval, ws, trust ? 0.0 : prm.growth_tol,
                                  trust ? 0.0 : pivot_floor_rel<T>() * jnorm, nullptr, opaque_lane(lane), &min_ratio, 0,
                                  P.n_rs);
    tick(t_fac);
    if (fail_col >= 0) { paused = 2; break; }
    solve(I.f, I.dx);
    for (int it = 0; it < prm.refine_steps; it++) {
      residual<true>(P, I.jv, I.f, I.dx, I.r, opaque_lane(lane));
      solve(I.r, I.r);
      for (int k = lane; k < m; k += kWave) I.dx[k] = I.dx[k] + I.r[k];
      WG_SYNC();
    }
    tick(t_sol);
    if (PAIR && lock_timeout) { step_failures++; break; }   // the solve vector never came free: the run ends
    int lbad;
    const double lmax = dx_max(I.dx, m, lane, kWave, &lbad);
    if (wave_any(lbad)) { step_failures++; break; }
    const double max_abs = wave_max(lmax);
    const double damp = max_abs > 0.4 ? 0.4 / max_abs : 1.0;
    // (dx in pivot-step order: each bus's unknowns found through ppos_k / qpos_k)
    apply_step(I, glob(P.ppos_k), glob(P.qpos_k), damp, n, lane, kWave);
    WG_SYNC();
    tick(t_upd);
    last_from = max_mis;
    iter++;
    steps++;
  }
  // final verdict (NewtonShell.AssembleResult precedence: NonFinite > Converged > Diverged > MaxIterations)
  int lnf = 0;
  for (int i0 = lane; i0 < n; i0 += 4 * kWave) {
    double a[4], b[4], c[4], e[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int i = i0 + u * kWave < n ? i0 + u * kWave : 0;
      a[u] = I.vm[i]; b[u] = I.va[i]; c[u] = I.p[i]; e[u] = I.q[i];
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      lnf |= !(isfinite(a[u]) && isfinite(b[u]) && isfinite(c[u] * P.base_mva) && isfinite(e[u] * P.base_mva));
  }
  const int nonfinite = wave_any(lnf) || !isfinite(max_mis);
  if (lane == 0) {
    st->iter = iter;
    st->start_mis = start_mis;
    st->max_mis = max_mis;
    st->min_ratio = min_ratio;
    st->step_failures = step_failures;
    st->fail_column = fail_col;
    st->last_from = last_from;
    st->engine_stop = lock_timeout ? GS_STOP_LOCK_TIMEOUT : 0;
    st->converged = converged;
    st->diverged = diverged;
    st->nonfinite = nonfinite;
    st->t_mismatch = t_mis;
    st->t_refactor = t_fac;
    st->t_solve = t_sol;
    st->t_update = t_upd;
    st->done = paused == 1 ? 0 : paused == 2 ? 2 : 1;
  }
}

// ------------------------------------------------------------------------------------------------ level-parallel
// The instances of a call spread over the whole GPU (schedule GS_SCHEDULE_LEVELS): each phase of a Newton step is its
// own kernel over a grid of (blocks of the phase, instances), and the refactor and the two triangular solves are one
// kernel per elimination level, a work-group per chunk of 8 columns (per pass of 64 rows), so the columns of a level
// run on as many compute units as the level has chunks. Kernel boundaries order the levels and make each level's
// factors visible to the next. The host captures one Newton step as a hipGraph and replays it; every kernel returns at
// once for an instance whose run has ended (its state's done flag). Reductions across the work-groups of an
// instance (max |f|, max |dx|, the smallest pivot ratio) go through atomics on the bit patterns of non-negative
// doubles. The arithmetic of every value is the batched kernel's: the same device functions over the same streams.

template <typename T>
__device__ __forceinline__ Inst<T> load_inst(const InstDesc* descs, const DevPlan& P, int b) {
  const InstDesc& d = descs[b];
  Inst<T> I;
  I.pspec = glob(d.pspec); I.qspec = glob(d.qspec); I.vm_start = glob(d.vm_start); I.va_start = glob(d.va_start);
  I.vm = glob(d.vm); I.va = glob(d.va); I.p = glob(d.p); I.q = glob(d.q);
  I.f = glob(d.f); I.dx = glob(d.dx); I.r = glob(d.r); I.jv = glob(d.jv);
  I.tpq = (GLOBAL v2d*)d.tpq;
  I.vals = glob(reinterpret_cast<T*>(d.vals));
  I.fac = I.vals + P.n_a;
  I.dval = glob(reinterpret_cast<T*>(d.dval));
  I.xg = glob(reinterpret_cast<T*>(d.xg));
  I.patch = glob(&descs[b].patch[0]);
  I.ygb = d.y_gb_own ? globv<v2d>(d.y_gb_own) : globv<v2d>(P.y_gb);
  I.n_patch = d.n_patch;
#pragma unroll
  for (int t = 0; t < kMaxPatch; t++) I.pidx[t] = t < I.n_patch ? d.patch[t].idx : -1;
  return I;
}

__device__ __forceinline__ void atomic_max_bits(uint64_t* p, double v) {  // v >= 0, or a NaN with its sign clear
  atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}
__device__ __forceinline__ double bits_double(uint64_t b) { return __longlong_as_double((long long)b); }

// Ends an instance's run (NewtonShell.AssembleResult precedence NonFinite > Converged > Diverged > MaxIterations):
// the non-finite scan over the buses by the calling work-group (nthreads threads, thread t), then the state, then the
// finished count. Called by one work-group of the instance.
__device__ void wide_finish(const DevPlan& P, InstState* st, const double* vm, const double* va, const double* p,
                            const double* q, double max_mis, int converged, int diverged, int t, int nthreads,
                            int* shared_flag, WideCtl* ctl) {
  int lnf = 0;
  for (int i = t; i < P.n; i += nthreads)
    lnf |= !(isfinite(vm[i]) && isfinite(va[i]) && isfinite(p[i] * P.base_mva) && isfinite(q[i] * P.base_mva));
  if (t == 0) *shared_flag = 0;
  __syncthreads();
  if (lnf) atomicOr(shared_flag, 1);
  __syncthreads();
  if (t == 0) {
    st->max_mis = max_mis;
    st->converged = converged;
    st->diverged = diverged;
    st->nonfinite = *shared_flag || !isfinite(max_mis);
    st->min_ratio = bits_double(st->ratio_bits);
    if (atomicCAS(&st->done, 0, 1) == 0) atomicAdd(&ctl->finished, 1);
  }
}

// start point and state of every instance (grid: blocks of 256 buses x instances)
__global__ __launch_bounds__(256) void wide_start(const InstDesc* __restrict__ descs, const DevPlan P, WideCtl* ctl) {
  const InstDesc& d = descs[blockIdx.y];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < P.n) { d.vm[i] = d.vm_start[i]; d.va[i] = d.va_start[i]; }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    InstState* st = d.state;
    st->iter = 0; st->done = 0; st->converged = 0; st->diverged = 0; st->nonfinite = 0; st->step_failures = 0;
    st->fail_column = -1; st->engine_stop = 0;
    st->start_mis = __builtin_nan(""); st->max_mis = __builtin_nan(""); st->min_ratio = 1.0;
    st->last_from = __builtin_nan("");
    st->t_mismatch = st->t_refactor = st->t_solve = st->t_update = 0;
    st->mis_bits = 0; st->dx_bits = 0; st->ratio_bits = (uint64_t)__double_as_longlong(1.0); st->jbad = 0;
    st->jn_bits = 0; st->jnorm = 0.0;
    if (blockIdx.y == 0) ctl->finished = 0;
  }
}

template <typename T>
__global__ __launch_bounds__(kWave) void wide_cells(const InstDesc* __restrict__ descs, const DevPlan P,
                                                    const NrParams prm) {
  InstState* st = descs[blockIdx.y].state;
  if (st->done) return;
  const Inst<T> I = load_inst<T>(descs, P, blockIdx.y);
  double jm;
  const TermsGlobal tg{I.tpq};
  const int bad = prm.refine_steps > 0
                      ? mismatch_cells<T, true>(P, I, threadIdx.x, 0, P.nnz_y, blockIdx.x, gridDim.x, tg, &jm)
                      : mismatch_cells<T, false>(P, I, threadIdx.x, 0, P.nnz_y, blockIdx.x, gridDim.x, tg, &jm);
  const int anybad = wave_any(bad);
  jm = wave_max(jm);
  if (threadIdx.x == 0) {
    if (anybad) atomicOr(&st->jbad, 1);
    atomic_max_bits(&st->jn_bits, jm);
  }
}

template <typename T>
__global__ __launch_bounds__(kWave) void wide_buses(const InstDesc* __restrict__ descs, const DevPlan P,
                                                    const NrParams prm) {
  InstState* st = descs[blockIdx.y].state;
  if (st->done) return;
  const Inst<T> I = load_inst<T>(descs, P, blockIdx.y);
  int bad;
  double jm;
  const double lmax = wave_max(prm.refine_steps > 0
                                   ? mismatch_buses<T, true>(P, I, threadIdx.x, 0, P.n, blockIdx.x, gridDim.x,
                                                             TermsGlobal{I.tpq}, &bad, &jm)
                                   : mismatch_buses<T, false>(P, I, threadIdx.x, 0, P.n, blockIdx.x, gridDim.x,
                                                              TermsGlobal{I.tpq}, &bad, &jm));
  const int anybad = wave_any(bad);
  jm = wave_max(jm);
  if (threadIdx.x == 0) {
    atomic_max_bits(&st->mis_bits, lmax);
    atomic_max_bits(&st->jn_bits, jm);
    if (anybad) atomicOr(&st->jbad, 1);
  }
}

// the verdicts after a mismatch evaluation (RunNewton :235-238, NewtonShell.DivergedBeyondLimit :177-178); one
// work-group per instance
__global__ __launch_bounds__(kWave) void wide_verdict(const InstDesc* __restrict__ descs, const DevPlan P,
                                                      const NrParams prm, WideCtl* ctl) {
  __shared__ int flag;
  const InstDesc& d = descs[blockIdx.y];
  InstState* st = d.state;
  if (st->done) return;
  const double max_mis = bits_double(st->mis_bits);
  const int bad = st->jbad;
  const int iter = st->iter;
  double start_mis = st->start_mis;
  if (iter == 0) start_mis = max_mis;
  const double jnorm = bits_double(st->jn_bits);
  __syncthreads();
  if (threadIdx.x == 0) {
    st->mis_bits = 0; st->jbad = 0; st->dx_bits = 0; st->jn_bits = 0;
    st->jnorm = jnorm;
    st->start_mis = start_mis;
    st->max_mis = max_mis;
  }
  const int converged = max_mis < prm.tolerance;
  const int diverged = !converged && max_mis > nan_max(prm.diverge_limit, start_mis);
  const bool stop = converged || diverged || iter >= prm.max_iterations || bad || !isfinite(max_mis);
  if (stop) wide_finish(P, st, d.vm, d.va, d.p, d.q, max_mis, converged, diverged, threadIdx.x, kWave, &flag, ctl);
}

// Chunks [first_chunk, end_chunk): one elimination level with a work-group per chunk, or (a single work-group) a run
// of consecutive narrow levels walked in stream order, as the batched kernel walks them (the stream's no-op rows and
// level-end fences order the levels inside the run).
template <typename T>
__global__ __launch_bounds__(kWave) void wide_refactor(const InstDesc* __restrict__ descs, const DevPlan P,
                                                       const NrParams prm, WideCtl* ctl, int first_chunk,
                                                       int end_chunk) {
  extern __shared__ __align__(16) unsigned char smem[];
  InstState* st = descs[blockIdx.y].state;
  if (st->done) return;
  const Inst<T> I = load_inst<T>(descs, P, blockIdx.y);
  const int per = (end_chunk - first_chunk) / (int)gridDim.x;
  const int c = first_chunk + blockIdx.x * per;
  const int row0 = P.rs_chunk_row[c], row1 = P.rs_chunk_row[c + per];
  double ratio = 1.0;
  const int fail = refactor<T, false, true>(P, I.vals, I.dval, lds(reinterpret_cast<T*>(smem)),
                                            prm.growth_tol, pivot_floor_rel<T>() * st->jnorm, nullptr,
                                            opaque_lane(threadIdx.x), &ratio, row0, row1);
  if (threadIdx.x == 0) {
    if (ratio < 1.0) atomicMin(reinterpret_cast<unsigned long long*>(&st->ratio_bits),
                               (unsigned long long)__double_as_longlong(ratio));
    if (fail >= 0) {
      // a static pivot failed its growth check: the instance pauses for a fresh host symbolic pass
      st->fail_column = fail;
      st->min_ratio = bits_double(st->ratio_bits);
      if (atomicCAS(&st->done, 0, 2) == 0) atomicAdd(&ctl->finished, 1);
    }
  }
}

// x (the solve vector, in global memory) = the row-permuted right-hand side (f, or the refinement residual r)
template <typename T>
__global__ __launch_bounds__(256) void wide_permute(const InstDesc* __restrict__ descs, const DevPlan P, int use_r) {
  const InstDesc& d = descs[blockIdx.y];
  if (d.state->done) return;
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r < P.m) reinterpret_cast<T*>(d.xg)[P.pinv[r]] = (T)(use_r ? d.r : d.f)[r];
}

// Passes [first_pass, end_pass) of a triangular solve: one level with a work-group per pass, or (a single work-group)
// a run of consecutive narrow levels in stream order.
template <typename T, bool BACK>
__global__ __launch_bounds__(kWave) void wide_solve(const InstDesc* __restrict__ descs, const DevPlan P,
                                                    int first_pass, int end_pass) {
  const InstDesc& d = descs[blockIdx.y];
  if (d.state->done) return;
  const int per = (end_pass - first_pass) / (int)gridDim.x;
  const int p = first_pass + blockIdx.x * per;
  const int* pass_step = BACK ? P.bs_pass_step : P.fs_pass_step;
  GLOBAL T* x = glob(reinterpret_cast<T*>(d.xg));
  extern __shared__ __align__(16) unsigned char smem[];
  solve_stream<T, BACK, false, GLOBAL T*, true>(globv<v2i>(BACK ? P.bs_rec : P.fs_rec), pass_step[p], pass_step[p + per],
                               glob(reinterpret_cast<const T*>(d.vals)) + P.n_a,
                               glob(reinterpret_cast<const T*>(d.dval)), x, lds(reinterpret_cast<T*>(smem)),
                               opaque_lane(threadIdx.x));
}

// dx (or r) = the column-permuted solve vector
template <typename T>
__global__ __launch_bounds__(256) void wide_unpermute(const InstDesc* __restrict__ descs, const DevPlan P,
                                                      int to_r) {
  const InstDesc& d = descs[blockIdx.y];
  if (d.state->done) return;
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < P.m) (to_r ? d.r : d.dx)[P.q[k]] = (double)reinterpret_cast<const T*>(d.xg)[k];
}

// r = f - J dx over 64 rows per work-group
__global__ __launch_bounds__(kWave) void wide_residual(const InstDesc* __restrict__ descs, const DevPlan P) {
  const InstDesc& d = descs[blockIdx.y];
  if (d.state->done) return;
  const int i = blockIdx.x * kWave + threadIdx.x;
  if (i >= P.m) return;
  GLOBAL const int* j_rp = glob(P.j_rp);
  GLOBAL const int* j_ci = glob(P.j_ci);
  GLOBAL const double* jv = glob(d.jv);
  GLOBAL const double* dx = glob(d.dx);
  double acc = 0.0;
  for (int k = j_rp[i]; k < j_rp[i + 1]; k++) acc += jv[k] * dx[j_ci[k]];
  d.r[i] = d.f[i] - acc;
}

__global__ __launch_bounds__(256) void wide_add_correction(const InstDesc* __restrict__ descs, const DevPlan P) {
  const InstDesc& d = descs[blockIdx.y];
  if (d.state->done) return;
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < P.m) d.dx[k] = d.dx[k] + d.r[k];
}

// max |dx| of the instance (NaN and inf order above every finite value)
__global__ __launch_bounds__(256) void wide_dx_max(const InstDesc* __restrict__ descs, const DevPlan P) {
  const InstDesc& d = descs[blockIdx.y];
  InstState* st = d.state;
  if (st->done) return;
  const int k = blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (k < P.m) {
    const double x = d.dx[k];
    v = isfinite(x) ? fabs(x) : __builtin_inf();
  }
  v = wave_max(v);
  if ((threadIdx.x % kWave) == 0) atomic_max_bits(&st->dx_bits, v);
}

// the damped update (NewtonShell.ApplyStep :128-136), or the end of the run when a step is not finite
__global__ __launch_bounds__(256) void wide_update(const InstDesc* __restrict__ descs, const DevPlan P,
                                                   WideCtl* ctl) {
  __shared__ int flag;
  const InstDesc& d = descs[blockIdx.y];
  InstState* st = d.state;
  if (st->done) return;
  const double max_abs = bits_double(st->dx_bits);
  if (!isfinite(max_abs)) {
    // NewtonShell.SolveStep: a non-finite step ends the run
    if (blockIdx.x == 0) {
      if (threadIdx.x == 0) st->step_failures++;
      wide_finish(P, st, d.vm, d.va, d.p, d.q, st->max_mis, 0, 0, threadIdx.x, 256, &flag, ctl);
    }
    return;
  }
  const double damp = max_abs > 0.4 ? 0.4 / max_abs : 1.0;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < P.n) {
    const int pp = P.ppos[i], qq = P.qpos[i];
    if (pp >= 0) d.va[i] = d.va[i] + damp * d.dx[pp];
    if (qq >= 0) {
      double v = d.vm[i] + damp * d.dx[qq];
      d.vm[i] = v < 0.01 ? 0.01 : v;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->last_from = st->max_mis;   // the mismatch this step started from
    st->iter = st->iter + 1;
  }
}

// ------------------------------------------------------------------------------------------------ work-group per instance
// Schedule GS_SCHEDULE_GROUP: the W wavefronts of one work-group run one instance's whole Newton loop, for networks
// whose elimination levels are wide enough to keep several waves busy but too many to pay a kernel each. Per step:
// the Y cells and the buses in blocks of 64 spread over the waves (a work-group barrier between the two passes, which
// make every injection term visible), the reductions of max |f|, max |J| and the Jacobian's finiteness through LDS;
// then the refactor and both triangular solves by the plan's segments: a level of several chunks (passes) split into W
// contiguous ranges, one per wave, each wave on its own LDS workspace, or a run of narrow levels walked by wave 0 in
// stream order; a barrier after each segment publishes its factors (solve values) to the next. The permutations, the
// refinement residual, max |dx| and the update spread over all threads. Every value is the batched kernel's: the same
// device functions over the same streams, each reduction over the same set of values.
constexpr int kGroupMaxWaves = 8;

// The group's reductions: each wave's (wave-uniform) values into one of two LDS slot sets, a barrier, every thread
// combines the W entries. Consecutive reductions alternate the sets, so a set is only rewritten after every wave has
// passed the barrier that follows its last read.
struct GroupReduce {
  double a[2][kGroupMaxWaves], b[2][kGroupMaxWaves];
  int c[2][kGroupMaxWaves];
  int fail[3];   // the refactor segments' failure keys (three in rotation: see group_refactor)
};

__device__ __forceinline__ void group_reduce(GroupReduce& g, int& set, int wave, int lane, int W, double* a, double* b,
                                             int* c, bool a_is_min) {
  if (lane == 0) { g.a[set][wave] = *a; g.b[set][wave] = *b; g.c[set][wave] = *c; }
  __syncthreads();
  double ra = g.a[set][0], rb = g.b[set][0];
  int rc = g.c[set][0];
  for (int w = 1; w < W; w++) {
    const double va = g.a[set][w];
    ra = a_is_min ? (va < ra ? va : ra) : nan_max(ra, va);
    rb = nan_max(rb, g.b[set][w]);
    rc |= g.c[set][w];
  }
  *a = ra; *b = rb; *c = rc;
  set ^= 1;
}

// The first and end chunk (pass) of this wave's share of segment sg.
__device__ __forceinline__ void group_share(v4i sg, int wave, int W, int* first, int* end) {
  if (sg.z) {
    const int c = sg.y - sg.x;
    *first = sg.x + (c * wave) / W;
    *end = sg.x + (c * (wave + 1)) / W;
  } else {
    *first = sg.x;
    *end = wave == 0 ? sg.y : sg.x;
  }
}

// Refactor by segments. Returns the failing pivot step, or -1. A failure ends the refactor after its segment; the
// failure reported is the first in stream order, as one wave walking the whole stream finds it (the waves' shares are
// in stream order, so the lowest failing wave's first). The failure keys rotate over three slots: slot g % 3 is
// written during segment g and read after its barrier, and thread 0 resets slot (g + 1) % 3 during segment g, after
// every wave has read it (segment g - 2) and before any wave writes it (segment g + 1).
template <typename T>
__device__ __forceinline__ int group_refactor(const DevPlan& P, const Inst<T>& I, LDSP T* ws, GroupReduce& g,
                                              double growth_tol, double abs_floor, double* min_ratio, int tid, int lane,
                                              int wave, int W) {
  GLOBAL const v4i* seg = globv<v4i>(P.gseg);
  GLOBAL const int* chunk_row = glob(P.rs_chunk_row);
  constexpr int kNone = 0x7FFFFFFF;
  if (tid == 0) { g.fail[0] = kNone; g.fail[1] = kNone; }
  __syncthreads();
  int failed = -1;
  for (int k = 0; k < P.n_gseg_rs; k++) {
    if (tid == 0) g.fail[(k + 1) % 3] = kNone;
    int c0, c1;
    group_share(seg[k], wave, W, &c0, &c1);
    if (c1 > c0) {
      const int f = refactor<T, false, true>(P, I.vals, I.dval, ws, growth_tol, abs_floor, nullptr,
                                             lane, min_ratio, chunk_row[c0], chunk_row[c1]);
      if (f >= 0 && lane == 0) atomicMin(&g.fail[k % 3], (wave << 24) | f);
    }
    __syncthreads();
    const int key = g.fail[k % 3];
    if (key != kNone) { failed = key & 0xFFFFFF; break; }
  }
  return failed;
}

// x = A^{-1} (permuted rhs), out = the column-permuted x: the forward then back segments.
template <typename T, bool XLDS, typename XP>
__device__ __forceinline__ void group_solve(const DevPlan& P, const Inst<T>& I, GLOBAL const double* rhs,
                                            GLOBAL double* out, XP x, LDSP T* scr, int tid, int nt, int lane,
                                            int wave, int W) {
  GLOBAL const int* pinv = glob(P.pinv);
  GLOBAL const int* q = glob(P.q);
  GLOBAL const v4i* seg = globv<v4i>(P.gseg);
  permute_in<XP, T>(x, pinv, rhs, P.m, tid, nt);
  __syncthreads();
  for (int k = 0; k < P.n_gseg_fs; k++) {
    int p0, p1;
    group_share(seg[P.n_gseg_rs + k], wave, W, &p0, &p1);
    if (p1 > p0)
      solve_stream<T, false, XLDS>(globv<v2i>(P.fs_rec), P.fs_pass_step[p0], P.fs_pass_step[p1], I.fac, I.dval, x,
                                   scr, lane);
    __syncthreads();
  }
  for (int k = 0; k < P.n_gseg_bs; k++) {
    int p0, p1;
    group_share(seg[P.n_gseg_rs + P.n_gseg_fs + k], wave, W, &p0, &p1);
    if (p1 > p0)
      solve_stream<T, true, XLDS>(globv<v2i>(P.bs_rec), P.bs_pass_step[p0], P.bs_pass_step[p1], I.fac, I.dval, x,
                                  scr, lane);
    __syncthreads();
  }
  for (int k = tid; k < P.m; k += nt) out[q[k]] = (double)x[k];
  __syncthreads();
}

template <typename T, bool XLDS>
__global__ __launch_bounds__(kWave * kGroupMaxWaves) void group_kernel(const InstDesc* __restrict__ descs,
                                                                        const DevPlan P, const NrParams prm) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ GroupReduce g;
  const int W = (int)blockDim.x / kWave, nt = (int)blockDim.x;
  const int tid = (int)threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  // LDS: each wave's refactor workspace (also its solve scratch), then (when it fits) the solve vector
  LDSP T* ws = lds(reinterpret_cast<T*>(smem)) + wave * ws_slots(P);
  LDSP T* xl = lds(reinterpret_cast<T*>(smem)) + W * ws_slots(P);
  const Inst<T> I = load_inst<T>(descs, P, blockIdx.x);
  InstState* st = descs[blockIdx.x].state;
  const int n = P.n, m = P.m;
  int set = 0;

  // (as the wavefront kernel: a fresh instance starts from its start point, any other finished one returns)
  const int was = prm.init ? 0 : st->done;
  const bool fresh = prm.init || was == kDoneFresh;
  if (fresh || prm.load_start) {
    for (int i = tid; i < n; i += nt) { I.vm[i] = I.vm_start[i]; I.va[i] = I.va_start[i]; }
    __syncthreads();
  }
  if (!fresh && was) return;
  int iter = fresh ? 0 : st->iter;
  double start_mis = fresh ? __builtin_nan("") : st->start_mis;
  double min_ratio = fresh ? 1.0 : st->min_ratio;
  double last_from = fresh ? __builtin_nan("") : st->last_from;
  int step_failures = fresh ? 0 : st->step_failures;
  int converged = 0, diverged = 0, paused = 0, fail_col = -1, steps = 0;
  double max_mis = __builtin_nan("");
  uint64_t t_mis = fresh ? 0 : st->t_mismatch, t_fac = fresh ? 0 : st->t_refactor;
  uint64_t t_sol = fresh ? 0 : st->t_solve, t_upd = fresh ? 0 : st->t_update;
  uint64_t clk = wall_clock64();
  auto tick = [&](uint64_t& acc) {
    const uint64_t now = wall_clock64();
    acc += now - clk;
    clk = now;
  };
  auto solve = [&](GLOBAL const double* rhs, GLOBAL double* out) {
    if constexpr (XLDS)
      group_solve<T, true>(P, I, rhs, out, xl, ws, tid, nt, opaque_lane(lane), wave, W);
    else
      group_solve<T, false>(P, I, rhs, out, I.xg, ws, tid, nt, opaque_lane(lane), wave, W);
  };
  for (;;) {
    if (steps >= prm.steps_this_launch) { paused = 1; break; }
    double jc;
    int bad;
    double jnorm;
    {
      // blocked: each wave its own blocks of whole rows and its own scratch (no barrier between the halves)
      LDSP double* mscr = reinterpret_cast<LDSP double*>(ws);
      max_mis = wave_max(prm.refine_steps > 0
                             ? mismatch_blocked<T, true>(P, I, opaque_lane(lane), wave, W, mscr, &bad, &jc)
                             : mismatch_blocked<T, false>(P, I, opaque_lane(lane), wave, W, mscr, &bad, &jc));
      bad = wave_any(bad);
      jnorm = wave_max(jc);
    }
    group_reduce(g, set, wave, lane, W, &jnorm, &max_mis, &bad, false);   // (also publishes f, p, q, J)
    tick(t_mis);
    if (iter == 0) start_mis = max_mis;
    if (max_mis < prm.tolerance) { converged = 1; break; }
    if (max_mis > nan_max(prm.diverge_limit, start_mis)) { diverged = 1; break; }
    if (iter >= prm.max_iterations) break;
    if (bad || !isfinite(max_mis)) break;  // NewtonShell.SolveStep refuses a non-finite J or f
    // (a plan the host made at this very state is the factorisation's own: its first step is not checked)
    const bool trust = prm.trust_first && steps == 0;
    double ratio = min_ratio;
    fail_col = group_refactor<T>(P, I, ws, g, trust ? 0.0 : prm.growth_tol,
                                 trust ? 0.0 : pivot_floor_rel<T>() * jnorm, &ratio, tid, opaque_lane(lane), wave, W);
    {
      double unused = 0.0;
      int none = 0;
      group_reduce(g, set, wave, lane, W, &ratio, &unused, &none, true);
    }
    min_ratio = ratio;
    tick(t_fac);
    if (fail_col >= 0) { paused = 2; break; }
    solve(I.f, I.dx);
    for (int it = 0; it < prm.refine_steps; it++) {
      residual(P, I.jv, I.f, I.dx, I.r, tid, nt);
      __syncthreads();
      solve(I.r, I.r);
      for (int k = tid; k < m; k += nt) I.dx[k] = I.dx[k] + I.r[k];
      __syncthreads();
    }
    tick(t_sol);
    double unused = 0.0;
    int lbad;
    double lmax = dx_max(I.dx, m, tid, nt, &lbad);
    lbad = wave_any(lbad);
    lmax = wave_max(lmax);
    group_reduce(g, set, wave, lane, W, &unused, &lmax, &lbad, false);
    if (lbad) { step_failures++; break; }
    const double damp = lmax > 0.4 ? 0.4 / lmax : 1.0;
    apply_step(I, glob(P.ppos), glob(P.qpos), damp, n, tid, nt);
    __syncthreads();
    tick(t_upd);
    last_from = max_mis;
    iter++;
    steps++;
  }
  // final verdict (NewtonShell.AssembleResult precedence: NonFinite > Converged > Diverged > MaxIterations)
  int lnf = 0;
  for (int i = tid; i < n; i += nt)
    lnf |= !(isfinite(I.vm[i]) && isfinite(I.va[i]) && isfinite(I.p[i] * P.base_mva) &&
             isfinite(I.q[i] * P.base_mva));
  lnf = wave_any(lnf);
  {
    double u0 = 0.0, u1 = 0.0;
    group_reduce(g, set, wave, lane, W, &u0, &u1, &lnf, false);
  }
  const int nonfinite = lnf || !isfinite(max_mis);
  if (tid == 0) {
    st->iter = iter;
    st->start_mis = start_mis;
    st->max_mis = max_mis;
    st->min_ratio = min_ratio;
    st->step_failures = step_failures;
    st->fail_column = fail_col;
    st->last_from = last_from;
    st->engine_stop = 0;
    st->converged = converged;
    st->diverged = diverged;
    st->nonfinite = nonfinite;
    st->t_mismatch = t_mis;
    st->t_refactor = t_fac;
    st->t_solve = t_sol;
    st->t_update = t_upd;
    st->done = paused == 1 ? 0 : paused == 2 ? 2 : 1;
  }
}

// ------------------------------------------------------------------------------------------------ iterative step
// NewtonShell.SolveStep's last resort (NewtonShell.cs:97-105): when the sparse LU refuses the Jacobian (the host's
// fresh symbolic pass found it singular), the step is solved by BiCGSTAB with an ILU(0) preconditioner on the FP64
// Jacobian (Sparse.cs:230-405: Ilu0 with the row-relative pivot guard, BiCGSTAB to min(1e-10, tol 1e-3) within
// 4 m + 200 iterations, its breakdown rules and the true-residual check). One Newton iteration of one instance: the
// mismatch and Jacobian (CSR, FP64) by the wave, the verdicts, then the factorisation, the Krylov loop and the damped
// update by one lane in the reference's sequential order. A rare path (an island with nothing to balance, such as a
// synchronous condenser cut off by its only branch), so one lane is enough. work: lu [nnz_j], then r, rhat, p, v, s,
// t, ph, sh [m each]; iw: [m].
__device__ void ilu0_apply_seq(const DevPlan& P, const double* lu, const double* r, double* z) {
  for (int i = 0; i < P.m; i++) {
    double s = r[i];
    for (int k = P.j_rp[i]; k < P.j_diag[i]; k++) s -= lu[k] * z[P.j_ci[k]];
    z[i] = s;
  }
  for (int i = P.m - 1; i >= 0; i--) {
    double s = z[i];
    for (int k = P.j_diag[i] + 1; k < P.j_rp[i + 1]; k++) s -= lu[k] * z[P.j_ci[k]];
    z[i] = s / lu[P.j_diag[i]];
  }
}

__device__ void csr_mul_seq(const DevPlan& P, const double* a, const double* x, double* y) {
  for (int i = 0; i < P.m; i++) {
    double s = 0;
    for (int k = P.j_rp[i]; k < P.j_rp[i + 1]; k++) s += a[k] * x[P.j_ci[k]];
    y[i] = s;
  }
}

__device__ double dot_seq(int n, const double* a, const double* b) {
  double s = 0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

// Returns whether BiCGSTAB produced x (A x = b to the tolerance).
__device__ bool ilu0_bicgstab(const DevPlan& P, const double* a, const double* b, double* x, double tol, double* work,
                              int* iw) {
  const int n = P.m;
  double* lu = work;
  double *r = lu + P.nnz_j, *rhat = r + n, *p = rhat + n, *v = p + n, *s = v + n, *t = s + n, *ph = t + n, *sh = ph + n;
  // ILU(0) (Sparse.cs:230-300)
  for (int k = 0; k < P.nnz_j; k++) lu[k] = a[k];
  for (int i = 0; i < n; i++) iw[i] = -1;
  for (int i = 0; i < n; i++) {
    for (int k = P.j_rp[i]; k < P.j_rp[i + 1]; k++) iw[P.j_ci[k]] = k;
    for (int k = P.j_rp[i]; k < P.j_diag[i]; k++) {
      const int j = P.j_ci[k];
      const double piv = lu[P.j_diag[j]];
      const double lij = (piv != 0.0) ? lu[k] / piv : 0.0;
      lu[k] = lij;
      if (lij == 0.0) continue;
      for (int kk = P.j_diag[j] + 1; kk < P.j_rp[j + 1]; kk++) {
        const int pos = iw[P.j_ci[kk]];
        if (pos != -1) lu[pos] -= lij * lu[kk];
      }
    }
    double row_scale = 0.0;
    for (int k = P.j_rp[i]; k < P.j_rp[i + 1]; k++) { const double av = fabs(a[k]); if (av > row_scale) row_scale = av; }
    const double guard = 1e-12 * (row_scale > 0.0 ? row_scale : 1.0);
    const double d = lu[P.j_diag[i]];
    if (fabs(d) < guard) lu[P.j_diag[i]] = d < 0 ? -guard : guard;
    for (int k = P.j_rp[i]; k < P.j_rp[i + 1]; k++) iw[P.j_ci[k]] = -1;
  }
  // BiCGSTAB (Sparse.cs:302-405)
  for (int i = 0; i < n; i++) { x[i] = 0.0; r[i] = b[i]; rhat[i] = b[i]; p[i] = 0; v[i] = 0; }
  const double bnorm = sqrt(dot_seq(n, b, b));
  if (bnorm == 0) {
    for (int i = 0; i < n; i++) if (b[i] != 0.0) return false;
    return true;
  }
  auto true_residual = [&]() {
    csr_mul_seq(P, a, x, t);
    double sum = 0.0;
    for (int i = 0; i < n; i++) { const double d = b[i] - t[i]; sum += d * d; }
    return sqrt(sum) / bnorm < tol;
  };
  double rho = 1, alpha = 1, omega = 1;
  const int max_iter = 4 * n + 200;
  for (int it = 0; it < max_iter; it++) {
    const double rho_new = dot_seq(n, rhat, r);
    if (fabs(rho_new) < 1e-300 || isnan(rho_new)) return false;
    const double beta = (rho_new / rho) * (alpha / omega);
    for (int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    ilu0_apply_seq(P, lu, p, ph);
    csr_mul_seq(P, a, ph, v);
    const double rhat_v = dot_seq(n, rhat, v);
    if (rhat_v == 0.0 || isnan(rhat_v)) return false;
    alpha = rho_new / rhat_v;
    for (int i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
    if (sqrt(dot_seq(n, s, s)) / bnorm < tol) {
      for (int i = 0; i < n; i++) x[i] += alpha * ph[i];
      return true_residual();
    }
    ilu0_apply_seq(P, lu, s, sh);
    csr_mul_seq(P, a, sh, t);
    const double tt = dot_seq(n, t, t);
    omega = tt > 0 ? dot_seq(n, t, s) / tt : 0.0;
    for (int i = 0; i < n; i++) { x[i] += alpha * ph[i] + omega * sh[i]; r[i] = s[i] - omega * t[i]; }
    rho = rho_new;
    if (sqrt(dot_seq(n, r, r)) / bnorm < tol) return true_residual();
    if (omega == 0) return false;
  }
  return false;
}

__global__ __launch_bounds__(kWave) void iterative_step_kernel(const InstDesc* __restrict__ descs, const DevPlan P,
                                                               const NrParams prm, double* work, int* iw) {
  const int lane = threadIdx.x;
  const Inst<double> I = load_inst<double>(descs, P, 0);
  InstState* st = descs[0].state;
  if (prm.load_start) {
    for (int i = lane; i < P.n; i += kWave) { I.vm[i] = I.vm_start[i]; I.va[i] = I.va_start[i]; }
    WG_SYNC();
  }
  int bad;
  double jnorm;
  // (the FP64 CSR copy only: this instance has no LU values)
  const double max_mis = mismatch_jacobian<double, true, false>(P, I, opaque_lane(lane), &bad, &jnorm);
  WG_SYNC();
  if (lane != 0) return;
  const int iter = st->iter;
  if (iter == 0) st->start_mis = max_mis;
  st->max_mis = max_mis;
  int finish = 0, converged = 0, diverged = 0;
  if (max_mis < prm.tolerance) { finish = 1; converged = 1; }
  else if (max_mis > nan_max(prm.diverge_limit, st->start_mis)) { finish = 1; diverged = 1; }
  else if (iter >= prm.max_iterations) finish = 1;
  else if (bad || !isfinite(max_mis)) finish = 1;
  if (!finish) {
    double* dx = descs[0].dx;
    st->step_failures++;   // the LU refused this step
    const double tol = fmin(1e-10, prm.tolerance * 1e-3);
    if (!ilu0_bicgstab(P, descs[0].jv, descs[0].f, dx, tol, work, iw)) {
      finish = 1;          // NewtonShell.SolveStep false: the run stops
    } else {
      double max_abs = 0;
      bool finite = true;
      for (int k = 0; k < P.m; k++) {
        if (!isfinite(dx[k])) { finite = false; break; }
        const double av = fabs(dx[k]);
        if (av > max_abs) max_abs = av;
      }
      if (!finite) {
        finish = 1;
      } else {
        const double damp = max_abs > 0.4 ? 0.4 / max_abs : 1.0;
        for (int i = 0; i < P.n; i++) {
          if (P.ppos[i] >= 0) I.va[i] = I.va[i] + damp * dx[P.ppos[i]];
          if (P.qpos[i] >= 0) {
            const double v = I.vm[i] + damp * dx[P.qpos[i]];
            I.vm[i] = v < 0.01 ? 0.01 : v;
          }
        }
        st->last_from = max_mis;
        st->iter = iter + 1;
      }
    }
  }
  if (finish) {
    int nf = !isfinite(max_mis);
    for (int i = 0; i < P.n; i++)
      nf |= !(isfinite(I.vm[i]) && isfinite(I.va[i]) && isfinite(I.p[i] * P.base_mva) && isfinite(I.q[i] * P.base_mva));
    st->converged = converged;
    st->diverged = diverged;
    st->nonfinite = nf;
    st->done = 1;
  } else {
    st->done = 0;
  }
}

// ------------------------------------------------------------------------------------------------ screening
// Line flows and the classification numbers of every instance's solved state (one work-group of 256 per instance):
// NewtonShell.ComputeLineFlows (NewtonShell.cs:699-734) per branch, the maximum loading and the overloads
// (ContingencyAnalysis.ScreenStateCore :1150-1194), Vmin / Vmax and the buses outside their limits (:1196-1234), and
// AssembleResult's losses as the sums of the bus injections in bus order (:649-690).
// Complex.Magnitude: the scaled hypotenuse |large| sqrt(1 + (small / large)^2).
__device__ __forceinline__ double cx_magnitude(double re, double im) {
  const double a = fabs(re), b = fabs(im);
  const double small = a < b ? a : b, large = a < b ? b : a;
  if (small == 0.0) return large;
  if (isinf(large) && !isnan(small)) return __builtin_inf();
  const double ratio = small / large;
  return large * sqrt(1.0 + ratio * ratio);
}

__global__ __launch_bounds__(256) void screen_kernel(const InstDesc* __restrict__ descs, const DevPlan P,
                                                     gs_screen_summary* out, double* loading, double* flows) {
  __shared__ double s_max[256];
  __shared__ int s_arg[256], s_over[256], s_under[256], s_overv[256];
  __shared__ double s_vmin[256], s_vmax[256];
  const InstDesc& d = descs[blockIdx.x];
  const int t = threadIdx.x;
  const double base = P.base_mva;
  double lmax = 0.0;
  int larg = -1, nover = 0;
  for (int k = t; k < P.nb; k += 256) {
    const int2 e = P.br_ends[k];
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    bool out = false;
    for (int t = 0; t < d.n_out; t++) out |= d.out[t] == k;
    if (e.x >= 0 && !out) {
      double sf, cf, st, ct;
      gs_fdlibm::sincos(d.va[e.x], &sf, &cf);
      gs_fdlibm::sincos(d.va[e.y], &st, &ct);
      const double vfr = d.vm[e.x] * cf, vfi = d.vm[e.x] * sf;  // DetMath.FromPolar
      const double vtr = d.vm[e.y] * ct, vti = d.vm[e.y] * st;
      const double* y = P.br_stamp + 8 * (size_t)k;  // Yff, Yft, Ytf, Ytt
      // I_from = Yff vf + Yft vt, I_to = Ytf vf + Ytt vt (Complex products (ac - bd) + (bc + ad) i)
      const double ifr = (y[0] * vfr - y[1] * vfi) + (y[2] * vtr - y[3] * vti);
      const double ifi = (y[1] * vfr + y[0] * vfi) + (y[3] * vtr + y[2] * vti);
      const double itr = (y[4] * vfr - y[5] * vfi) + (y[6] * vtr - y[7] * vti);
      const double iti = (y[5] * vfr + y[4] * vfi) + (y[7] * vtr + y[6] * vti);
      // S = V conj(I) * base
      const double sfr = (vfr * ifr - vfi * -ifi) * base, sfi = (vfi * ifr + vfr * -ifi) * base;
      const double str = (vtr * itr - vti * -iti) * base, sti = (vti * itr + vtr * -iti) * base;
      const double mf = cx_magnitude(sfr, sfi), mt = cx_magnitude(str, sti);
      const double mva = (isnan(mf) || isnan(mt)) ? __builtin_nan("") : (mf > mt ? mf : mt);
      const double rate = P.br_rate[k];
      v[0] = sfr; v[1] = sfi; v[2] = str; v[3] = sti;
      v[4] = rate > 0 ? mva / rate * 100.0 : 0.0;
    }
    if (loading) loading[(size_t)blockIdx.x * P.nb + k] = v[4];
    if (flows) {
      double* fo = flows + ((size_t)blockIdx.x * P.nb + k) * 4;
      fo[0] = v[0]; fo[1] = v[1]; fo[2] = v[2]; fo[3] = v[3];
    }
    if (v[4] > lmax) { lmax = v[4]; larg = k; }
    nover += v[4] > 100.0;
  }
  double vmin = __builtin_inf(), vmax = -__builtin_inf();
  int nunder = 0, novv = 0;
  for (int i = t; i < P.n; i += 256) {
    const double vm = d.vm[i];
    vmin = vm < vmin ? vm : vmin;
    vmax = vm > vmax ? vm : vmax;
    const double lo = isnan(P.bus_lo[i]) ? 0.94 : P.bus_lo[i];
    const double hi = isnan(P.bus_hi[i]) ? 1.06 : P.bus_hi[i];
    if (vm < lo) nunder++;
    else if (vm > hi) novv++;
  }
  s_max[t] = lmax; s_arg[t] = larg; s_over[t] = nover; s_under[t] = nunder; s_overv[t] = novv;
  s_vmin[t] = vmin; s_vmax[t] = vmax;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      // the largest loading, the lowest branch index among equal ones
      const bool take = s_max[t + w] > s_max[t] || (s_max[t + w] == s_max[t] && s_arg[t + w] >= 0 &&
                                                    (s_arg[t] < 0 || s_arg[t + w] < s_arg[t]));
      if (take) { s_max[t] = s_max[t + w]; s_arg[t] = s_arg[t + w]; }
      s_over[t] += s_over[t + w]; s_under[t] += s_under[t + w]; s_overv[t] += s_overv[t + w];
      s_vmin[t] = s_vmin[t + w] < s_vmin[t] ? s_vmin[t + w] : s_vmin[t];
      s_vmax[t] = s_vmax[t + w] > s_vmax[t] ? s_vmax[t + w] : s_vmax[t];
    }
    __syncthreads();
  }
  if (t == 0) {
    double lp = 0.0, lq = 0.0;
    for (int i = 0; i < P.n; i++) { lp += d.p[i] * base; lq += d.q[i] * base; }
    gs_screen_summary o;
    o.max_loading_pct = s_max[0];
    o.max_loading_branch = s_max[0] > 0.0 ? s_arg[0] : -1;
    o.overloads = s_over[0];
    o.vmin = s_vmin[0];
    o.vmax = s_vmax[0];
    o.undervoltage = s_under[0];
    o.overvoltage = s_overv[0];
    o.loss_mw = lp;
    o.loss_mvar = lq;
    out[blockIdx.x] = o;
  }
}

// ------------------------------------------------------------------------------------------------ seam solves
// One CSR system on a plan made from that system's own factorisation: the numeric refactor replays the reference's
// arithmetic (SparseLuFactorization.Refactor), refusing exactly where it would (a pivot that moves, or one at or below
// its row floor), then the reference's solve.
__global__ __launch_bounds__(kWave) void seam_solve_kernel(const SeamDesc* __restrict__ descs, const DevPlan P) {
  extern __shared__ __align__(16) unsigned char smem[];
  const SeamDesc& d = descs[blockIdx.x];
  const int lane = threadIdx.x, m = P.m;
  LDSP double* ws = lds(reinterpret_cast<double*>(smem));
  GLOBAL const double* a = glob(d.a);
  GLOBAL double* rowfloor = glob(d.rowfloor);
  GLOBAL double* vals = glob(d.vals);
  GLOBAL const int* j_rp = glob(P.j_rp);
  GLOBAL const int* ch_a_src = glob(P.ch_a_src);
  // row floors of the matrix as given (1e-12 x the largest |entry| of each row)
  for (int i = lane; i < m; i += kWave) {
    double mx = 0.0;
    for (int k = j_rp[i]; k < j_rp[i + 1]; k++) mx = fabs(a[k]) > mx ? fabs(a[k]) : mx;
    rowfloor[i] = mx * 1e-12;
  }
  for (int e = lane; e < P.n_a; e += kWave) vals[e] = a[ch_a_src[e]];
  WG_SYNC();
  double ratio = 1.0;
  const int fail =
      refactor<double, true>(P, vals, glob(d.dval), ws, 0.0, 0.0, rowfloor, lane, &ratio, 0, P.n_rs);
  if (fail >= 0) {
    if (lane == 0) *d.status = 1;
    return;
  }
  lu_solve<double, false>(P, vals + P.n_a, glob(d.dval), glob(d.b), glob(d.x), glob(d.xs), ws, lane);
  if (lane == 0) *d.status = 0;
}

// ------------------------------------------------------------------------------------------------ small kernels
__global__ void sincos_kernel(const double* x, double* s, double* c, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) gs_fdlibm::sincos(x[i], &s[i], &c[i]);
}

__global__ void fp64_probe_kernel(double* out, double a) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    out[0] = (1.0 + a) - 1.0;
    double s, c;
    gs_fdlibm::sincos(3.141592653589793 / 7.0, &s, &c);
    out[1] = s;
    out[2] = c;
  }
}

// host-side launchers (engine.cpp)
hipError_t launch_nr_batch(int lu_fp32, int x_in_lds, int pair, const InstDesc* descs, const DevPlan& plan,
                           int count, const NrParams& prm0, size_t lds_bytes, hipStream_t stream) {
  NrParams prm = prm0;
  prm.n_inst = count;
  if (pair && !x_in_lds) return hipErrorInvalidValue;
  if ((lu_fp32 ? plan.n_mb4 : plan.n_mb) <= 0) return hipErrorInvalidValue;   // (the blocked mismatch pass's blocks)
  const dim3 grid(pair ? (count + kPairWaves - 1) / kPairWaves : count), block(pair ? kWave * kPairWaves : kWave);
  if (lu_fp32) {
    if (pair)
      hipLaunchKernelGGL((nr_batch_kernel<float, true, true>), grid, block, lds_bytes, stream, descs, plan, prm);
    else if (x_in_lds)
      hipLaunchKernelGGL((nr_batch_kernel<float, true, false>), grid, block, lds_bytes, stream, descs, plan, prm);
    else
      hipLaunchKernelGGL((nr_batch_kernel<float, false, false>), grid, block, lds_bytes, stream, descs, plan, prm);
  } else {
    if (pair)
      hipLaunchKernelGGL((nr_batch_kernel<double, true, true>), grid, block, lds_bytes, stream, descs, plan, prm);
    else if (x_in_lds)
      hipLaunchKernelGGL((nr_batch_kernel<double, true, false>), grid, block, lds_bytes, stream, descs, plan, prm);
    else
      hipLaunchKernelGGL((nr_batch_kernel<double, false, false>), grid, block, lds_bytes, stream, descs, plan, prm);
  }
  return hipGetLastError();
}

hipError_t launch_group(int lu_fp32, int x_in_lds, int waves, const InstDesc* descs, const DevPlan& plan, int count,
                        const NrParams& prm0, size_t lds_bytes, hipStream_t stream) {
  NrParams prm = prm0;
  prm.n_inst = count;
  if (waves < 1 || waves > kGroupMaxWaves) return hipErrorInvalidValue;
  if ((lu_fp32 ? plan.n_mb4 : plan.n_mb) <= 0) return hipErrorInvalidValue;   // (the blocked mismatch pass's blocks)
  const dim3 grid(count), block(kWave * waves);
  if (lu_fp32) {
    if (x_in_lds)
      hipLaunchKernelGGL((group_kernel<float, true>), grid, block, lds_bytes, stream, descs, plan, prm);
    else
      hipLaunchKernelGGL((group_kernel<float, false>), grid, block, lds_bytes, stream, descs, plan, prm);
  } else {
    if (x_in_lds)
      hipLaunchKernelGGL((group_kernel<double, true>), grid, block, lds_bytes, stream, descs, plan, prm);
    else
      hipLaunchKernelGGL((group_kernel<double, false>), grid, block, lds_bytes, stream, descs, plan, prm);
  }
  return hipGetLastError();
}

hipError_t launch_wide_start(const InstDesc* descs, const DevPlan& plan, int count, WideCtl* ctl, hipStream_t stream) {
  hipLaunchKernelGGL(wide_start, dim3((plan.n + 255) / 256, count), dim3(256), 0, stream, descs, plan, ctl);
  return hipGetLastError();
}


// One Newton step of every instance, as the kernel sequence the host captures: mismatch, verdict, refactor level by
// level, the solves level by level (with the refinement steps), the update. level_chunks / fs_level_pass /
// bs_level_pass: first chunk (pass) of each level, and the end.
template <typename T>
static void enqueue_wide_step(const InstDesc* descs, const DevPlan& P, int count, const NrParams& prm, WideCtl* ctl,
                              const std::vector<int>& level_chunks, const std::vector<int>& fs_level_pass,
                              const std::vector<int>& bs_level_pass, size_t ws_bytes, const WideSplit& split,
                              hipStream_t s) {
  const dim3 w(kWave), b256(256);
  const unsigned nb64 = (unsigned)((P.nnz_y + kWave - 1) / kWave), nbus64 = (unsigned)((P.n + kWave - 1) / kWave);
  const unsigned m256 = (unsigned)((P.m + 255) / 256), m64 = (unsigned)((P.m + kWave - 1) / kWave);
  const unsigned n256 = (unsigned)((P.n + 255) / 256), cnt = (unsigned)count;
  hipLaunchKernelGGL(wide_cells<T>, dim3(nb64, cnt), w, 0, s, descs, P, prm);
  hipLaunchKernelGGL(wide_buses<T>, dim3(nbus64, cnt), w, 0, s, descs, P, prm);
  hipLaunchKernelGGL(wide_verdict, dim3(1, cnt), w, 0, s, descs, P, prm, ctl);
  // levels of at most kNarrow chunks (passes) run back to back on one work-group; wider levels one kernel each
  auto segments = [](const std::vector<int>& level_first, int narrow, auto&& launch) {
    int run0 = -1, run1 = -1;
    for (size_t l = 0; l + 1 < level_first.size(); l++) {
      const int a = level_first[l], b = level_first[l + 1];
      if (b <= a) continue;
      if (b - a <= narrow) {
        if (run0 < 0) run0 = a;
        run1 = b;
        continue;
      }
      if (run0 >= 0) { launch(run0, run1, 1); run0 = -1; }
      launch(a, b, b - a);
    }
    if (run0 >= 0) launch(run0, run1, 1);
  };
  auto slice = [](const std::vector<int>& v, int a, int b) {   // the table of levels [a, b)
    return std::vector<int>(v.begin() + a, v.begin() + b + 1);
  };
  auto refactor_levels = [&](const std::vector<int>& tab, hipStream_t st) {
    segments(tab, kNarrow, [&](int c0, int c1, int blocks) {
      hipLaunchKernelGGL(wide_refactor<T>, dim3(blocks, cnt), w, ws_bytes, st, descs, P, prm, ctl, c0, c1);
    });
  };
  auto forward_levels = [&](const std::vector<int>& tab) {
    segments(tab, kNarrowSolve, [&](int p0, int p1, int blocks) {
      hipLaunchKernelGGL((wide_solve<T, false>), dim3(blocks, cnt), w, ws_bytes, s, descs, P, p0, p1);
    });
  };
  const int nl = (int)level_chunks.size() - 1, nf = (int)fs_level_pass.size() - 1;
  const bool fork = split.side != nullptr && split.trunk_level > 0 && split.trunk_level < nl && split.fs_level > 0 &&
                    split.fs_level <= nf;
  if (fork) {
    // the trunk of the refactor on the side branch; the forward levels that read none of its columns meanwhile
    refactor_levels(slice(level_chunks, 0, split.trunk_level), s);
    (void)hipEventRecord(split.fork, s);
    (void)hipStreamWaitEvent(split.side, split.fork, 0);
    refactor_levels(slice(level_chunks, split.trunk_level, nl), split.side);
    (void)hipEventRecord(split.join, split.side);
    hipLaunchKernelGGL(wide_permute<T>, dim3(m256, cnt), b256, 0, s, descs, P, 0);
    forward_levels(slice(fs_level_pass, 0, split.fs_level));
    (void)hipStreamWaitEvent(s, split.join, 0);
  } else {
    refactor_levels(level_chunks, s);
  }
  auto solve = [&](int use_r, bool forward_from_split) {
    if (forward_from_split) {
      forward_levels(slice(fs_level_pass, split.fs_level, nf));
    } else {
      hipLaunchKernelGGL(wide_permute<T>, dim3(m256, cnt), b256, 0, s, descs, P, use_r);
      forward_levels(fs_level_pass);
    }
    segments(bs_level_pass, kNarrowSolve, [&](int p0, int p1, int blocks) {
      hipLaunchKernelGGL((wide_solve<T, true>), dim3(blocks, cnt), w, ws_bytes, s, descs, P, p0, p1);
    });
    hipLaunchKernelGGL(wide_unpermute<T>, dim3(m256, cnt), b256, 0, s, descs, P, use_r);
  };
  solve(0, fork);
  for (int it = 0; it < prm.refine_steps; it++) {
    hipLaunchKernelGGL(wide_residual, dim3(m64, cnt), w, 0, s, descs, P);
    solve(1, false);
    hipLaunchKernelGGL(wide_add_correction, dim3(m256, cnt), b256, 0, s, descs, P);
  }
  hipLaunchKernelGGL(wide_dx_max, dim3(m256, cnt), b256, 0, s, descs, P);
  hipLaunchKernelGGL(wide_update, dim3(n256, cnt), b256, 0, s, descs, P, ctl);
}

hipError_t launch_wide_step(int lu_fp32, const InstDesc* descs, const DevPlan& plan, int count, const NrParams& prm,
                            WideCtl* ctl, const std::vector<int>& level_chunks, const std::vector<int>& fs_level_pass,
                            const std::vector<int>& bs_level_pass, size_t ws_bytes, const WideSplit& split,
                            hipStream_t stream) {
  if (lu_fp32)
    enqueue_wide_step<float>(descs, plan, count, prm, ctl, level_chunks, fs_level_pass, bs_level_pass, ws_bytes,
                             split, stream);
  else
    enqueue_wide_step<double>(descs, plan, count, prm, ctl, level_chunks, fs_level_pass, bs_level_pass, ws_bytes,
                              split, stream);
  return hipGetLastError();
}

hipError_t launch_iterative_step(const InstDesc* descs, const DevPlan& plan, const NrParams& prm, double* work,
                                 int* iw, hipStream_t stream) {
  hipLaunchKernelGGL(iterative_step_kernel, dim3(1), dim3(kWave), 0, stream, descs, plan, prm, work, iw);
  return hipGetLastError();
}

hipError_t launch_screen(const InstDesc* descs, const DevPlan& plan, int count, gs_screen_summary* out,
                         double* loading, double* flows, hipStream_t stream) {
  hipLaunchKernelGGL(screen_kernel, dim3(count), dim3(256), 0, stream, descs, plan, out, loading, flows);
  return hipGetLastError();
}

hipError_t launch_seam(const SeamDesc* descs, const DevPlan& plan, int count, size_t lds_bytes, hipStream_t stream) {
  hipLaunchKernelGGL(seam_solve_kernel, dim3(count), dim3(kWave), lds_bytes, stream, descs, plan);
  return hipGetLastError();
}

hipError_t launch_sincos(const double* x, double* s, double* c, int64_t n, hipStream_t stream) {
  int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(sincos_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, x, s, c, n);
  return hipGetLastError();
}

hipError_t launch_fp64_probe(double* out, hipStream_t stream) {
  hipLaunchKernelGGL(fp64_probe_kernel, dim3(1), dim3(64), 0, stream, out, 9.094947017729282e-13);
  return hipGetLastError();
}

}  // namespace gs
