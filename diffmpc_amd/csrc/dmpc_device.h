// Device helpers shared by the fused kernels: DPP group broadcasts and
// reductions inside a wavefront, workgroup reductions, float64 dot products
// of a lane group against a matrix row held per lane.  Every function here
// must be called by all lanes of the wavefront (DPP and readlane read other
// lanes), i.e. outside lane-divergent branches.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace dmpc {
namespace {

// Sum over the G lanes of a block group, result in every lane of the group
// (DPP butterflies: quad swaps, then half-row and row mirrors).
template <int ctrl>
__device__ __forceinline__ double dpp_mov(double v) {
  long long x = __double_as_longlong(v);
  x = __builtin_amdgcn_mov_dpp(x, ctrl, 0xf, 0xf, true);
  return __longlong_as_double(x);
}
// Value of lane src (wave-uniform result, read as scalars).
__device__ __forceinline__ double lane_value(double v, int src) {
  const long long x = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)x, src);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(x >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// gfx950 row exchanges of a 64-bit value (v_permlane16_swap /
// v_permlane32_swap with both operands the same register, per 32-bit half):
// with rows r0..r3 of the wavefront,
//   swap16 -> {[r0, r0, r2, r2], [r1, r1, r3, r3]}
//   swap32 -> {[r0, r1, r0, r1], [r2, r3, r2, r3]}
__device__ __forceinline__ void rows_swap16(double v, double& even, double& odd) {
  const long long x = __double_as_longlong(v);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)x, (unsigned)x, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(x >> 32), (unsigned)(x >> 32), false, false);
  even = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
  odd = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}
__device__ __forceinline__ void rows_swap32(double v, double& low, double& high) {
  const long long x = __double_as_longlong(v);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)x, (unsigned)x, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(x >> 32), (unsigned)(x >> 32), false, false);
  low = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
  high = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}

// G = 64 is a whole wavefront: the four row sums are combined across rows
// with row exchanges as (R0 + R1) + (R2 + R3), in every lane.
template <int G>
__device__ __forceinline__ double gsum(double v) {
  static_assert(G == 4 || G == 8 || G == 16 || G == 64, "block groups of 4, 8, 16 or 64 lanes");
  v += dpp_mov<0xB1>(v);                        // quad_perm [1,0,3,2]
  v += dpp_mov<0x4E>(v);                        // quad_perm [2,3,0,1]
  if constexpr (G >= 8) v += dpp_mov<0x141>(v);  // row_half_mirror
  if constexpr (G >= 16) v += dpp_mov<0x140>(v); // row_mirror
  if constexpr (G == 64) {
    double e, o, lo, hi;
    rows_swap16(v, e, o);  // [R0, R0, R2, R2], [R1, R1, R3, R3]
    rows_swap32(e + o, lo, hi);
    v = lo + hi;
  }
  return v;
}


// Sum over a full wave: DPP sums inside each 16-lane row, then the four row
// sums are read as scalars.  The result is wave-uniform.
__device__ __forceinline__ double wave_sum(double v) {
  v = gsum<16>(v);
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}

__device__ __forceinline__ double nanmax(double a, double b) {
  if (a != a || b != b) return __longlong_as_double(0x7ff8000000000000ULL);
  return a > b ? a : b;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = nanmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Sum over the workgroup (at most 16 waves) through one barrier; every
// thread gets the same, bit-identical result.  After the barrier lane i of
// each 16-lane row reads wave i's sum (one LDS access per lane, all in
// parallel) and the row sums them with the DPP butterfly of gsum: pairwise
// additions whose operands are only swapped between lanes, so every lane
// ends with the same bits.
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const int nw = blockDim.x >> 6;
  const int k = threadIdx.x & 15;
  return gsum<16>(k < nw ? red[k] : 0.0);
}

// block_sum for workgroups of up to 1024 threads with `red` holding 64
// doubles: each 16-lane row publishes its DPP row sum (no cross-row
// combination inside the wave), and after the barrier lane k of every row
// adds the partials k, k + 16, k + 32, k + 48 and the row sums those with
// the butterfly.  Every thread reads the same partials in the same order, so
// every thread gets the same bits.  Fewer instructions than block_sum.
// `red` must hold zeros past the workgroup's rows (the caller clears all 64
// entries once; rows beyond blockDim.x / 16 are never written), so the four
// partial reads need no EXEC mask.
// The two halves are separate so that a caller can issue loads that only
// need the barrier (neighbour vector entries) between them.
// `row`: the publishing row's index in `red` (threadIdx.x / 16, or the
// row's index inside its team when a workgroup holds several instances)
__device__ __forceinline__ void block_sum_rows_publish(double v, double* red, int row) {
  v = gsum<16>(v);
  if ((threadIdx.x & 15) == 0) red[row] = v;
  __syncthreads();
}
__device__ __forceinline__ void block_sum_rows_publish(double v, double* red) {
  block_sum_rows_publish(v, red, threadIdx.x >> 4);
}
__device__ __forceinline__ void block_sum_rows_load(const double* red, double (&q)[4]) {
  const double* p = red + (threadIdx.x & 15);
#pragma unroll
  for (int m = 0; m < 4; ++m) q[m] = p[16 * m];
}
__device__ __forceinline__ double block_sum_rows_sum(const double (&q)[4]) {
  double s = 0.0;
#pragma unroll
  for (int m = 0; m < 4; ++m) s += q[m];
  return gsum<16>(s);
}
__device__ __forceinline__ double block_sum_rows_finish(const double* red) {
  double q[4];
  block_sum_rows_load(red, q);
  return block_sum_rows_sum(q);
}
__device__ __forceinline__ double block_sum_rows(double v, double* red, int row) {
  block_sum_rows_publish(v, red, row);
  return block_sum_rows_finish(red);
}
__device__ __forceinline__ double block_sum_rows(double v, double* red) {
  return block_sum_rows(v, red, threadIdx.x >> 4);
}

// a / b for the PCG step lengths: the hardware reciprocal refined by two
// Newton steps and one residual correction of the quotient (within an ulp of
// the correctly rounded quotient for normal operands; fewer instructions than
// the IEEE division sequence).
__device__ __forceinline__ double div_newton(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  r = fma(fma(-b, r, 1.0), r, r);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}

__device__ __forceinline__ double block_max(double v, double* red) {
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const int nw = blockDim.x >> 6;
  double s = red[0];
  for (int k = 1; k < nw; ++k) s = nanmax(s, red[k]);
  return s;
}

// block_max over the rows of one team (rows 0 .. nrows - 1 of `red`, `row`
// this row's index there).  Maxima are exact, so for values without signed
// zeros (the step norms it takes) the result equals block_max's.
__device__ __forceinline__ double block_max_rows(double v, double* red, int row, int nrows) {
#pragma unroll
  for (int off = 8; off > 0; off >>= 1) v = nanmax(v, __shfl_xor(v, off, 16));
  if ((threadIdx.x & 15) == 0) red[row] = v;
  __syncthreads();
  double s = red[0];
  for (int k = 1; k < nrows; ++k) s = nanmax(s, red[k]);
  return s;
}

__device__ __forceinline__ double nan_to_num(double v) {
  if (v != v) return 0.0;
  if (v == __longlong_as_double(0x7ff0000000000000ULL)) return 1.7976931348623157e308;
  if (v == __longlong_as_double(0xfff0000000000000ULL)) return -1.7976931348623157e308;
  return v;
}


// float -> double conversion that the optimizer cannot move out of the PCG
// loop: the asm takes the (uniform) iteration counter as an operand, so it is
// not loop-invariant, but it has no side effects and schedules freely.
// Without it the compiler hoists the conversions of the A/B registers and
// keeps the double copies, which do not fit in the register file.
__device__ __forceinline__ double cvt_in_loop(float f, int salt) {
  double d;
  asm("v_cvt_f64_f32 %0, %1" : "=v"(d) : "v"(f), "s"(salt));
  return d;
}
__device__ __forceinline__ double cvt_in_loop(double d, int) { return d; }

// `off`, marked as depending on `dep` (an empty asm, no instruction): loads
// addressed with the result cannot be issued before `dep` is computed.  In a
// sequential substitution this keeps the scheduler from hoisting every row's
// loads above the whole chain, which with one time block per lane group would
// hold more registers than the wave has and spill.
__device__ __forceinline__ int after(int off, double dep) {
  asm("" : "+v"(off) : "v"(dep));
  return off;
}

// ---------------------------------------------------------------------------
// Broadcast of lane n of each G-lane block group to the whole group with DPP
// (no LDS traffic): G = 16 is one DPP row (row_newbcast:n, one 64-bit move on
// gfx950); G = 4 a quad (quad_perm [n,n,n,n]); G = 8 is half a row (two
// bank-masked row_newbcast moves).
template <int n>
__device__ __forceinline__ double dpp_rowbcast(double v) {
  long long x = __double_as_longlong(v);
  x = __builtin_amdgcn_mov_dpp(x, 0x150 + n, 0xf, 0xf, true);
  return __longlong_as_double(x);
}
template <int n>
__device__ __forceinline__ double dpp_quadbcast(double v) {
  long long x = __double_as_longlong(v);
  x = __builtin_amdgcn_mov_dpp(x, n | (n << 2) | (n << 4) | (n << 6), 0xf, 0xf, true);
  return __longlong_as_double(x);
}

template <int G, int n>
__device__ __forceinline__ double gbcast(double v, int lane) {
  static_assert(n < G, "source lane outside the group");
  if constexpr (G == 64) {
    return lane_value(v, n);  // the group is the wavefront
  } else if constexpr (G == 16) {
    return dpp_rowbcast<n>(v);
  } else if constexpr (G == 4) {
    return dpp_quadbcast<n>(v);
  } else {
    static_assert(G == 8, "block groups of 4, 8, 16 or 64 lanes");
    // groups are 8-aligned in the workgroup, two per DPP row: broadcast lane
    // n into banks 0-1 (lanes 0-7) and lane 8 + n into banks 2-3 (lanes 8-15)
    // of the same register; the bank masks do the selection
    const long long x = __double_as_longlong(v);
    const long long lo = __builtin_amdgcn_mov_dpp(x, 0x150 + n, 0xf, 0x3, false);
    return __longlong_as_double(__builtin_amdgcn_update_dpp(lo, x, 0x150 + n + 8, 0xf, 0xc, false));
  }
}

// gbcast with a source lane known only after loop unrolling (a constant
// there, so the chain of tests folds away).
template <int G, int n = 0>
__device__ __forceinline__ double gbcast_at(double v, int src, int lane) {
  if constexpr (n == G - 1) {
    return gbcast<G, n>(v, lane);
  } else {
    return src == n ? gbcast<G, n>(v, lane) : gbcast_at<G, n + 1>(v, src, lane);
  }
}

// acc += (value of `src` in lane n of the group) * scale, n known after
// unrolling.  16-lane groups: one v_fmac_f64_dpp with row_newbcast:n (no
// broadcast temporary; the s_nop covers the DPP read-after-VALU-write wait);
// other groups: gbcast and an FMA.
template <int n>
__device__ __forceinline__ void dpp_fmac16(double& acc, double src, double scale) {
  asm("s_nop 1\n"
      "v_fmac_f64_dpp %0, %1, %2 row_newbcast:%c3 row_mask:0xf bank_mask:0xf"
      : "+v"(acc)
      : "v"(src), "v"(scale), "i"(n));
}
template <int G, int n = 0>
__device__ __forceinline__ void bcast_fmac_at(double& acc, double src, double scale, int from, int lane) {
  if constexpr (n == G - 1) {
    if constexpr (G == 16) {
      dpp_fmac16<n>(acc, src, scale);
    } else {
      acc = fma(gbcast<G, n>(src, lane), scale, acc);
    }
  } else {
    if (from == n) {
      if constexpr (G == 16) {
        dpp_fmac16<n>(acc, src, scale);
      } else {
        acc = fma(gbcast<G, n>(src, lane), scale, acc);
      }
    } else {
      bcast_fmac_at<G, n + 1>(acc, src, scale, from, lane);
    }
  }
}

// 64-lane groups (the group is the wavefront): DPP broadcasts stay inside a
// 16-lane row, and a readlane broadcast costs two v_readlane_b32 per term.
// Three row exchanges per 32-bit half give each row of v in every row, and
// each term is then one v_fmac_f64_dpp row_newbcast of the right copy, in
// the order j = 0..M-1 with one rounding per term, as gdot_chain.
__device__ __forceinline__ void rows_to_all(double v, double (&b)[4]) {
  double l, h;
  rows_swap32(v, l, h);     // [r0, r1, r0, r1], [r2, r3, r2, r3]
  rows_swap16(l, b[0], b[1]);
  rows_swap16(h, b[2], b[3]);
}
// acc + sum_{k<16} s[k] * (lane k of the row in b); the leading s_nop 4
// covers the DPP read-after-VALU-write wait states of b
__device__ __forceinline__ double fmac_row16(double acc, double b, const double (&s)[16]) {
#define DMPC_FMAC16(op, k) "v_fmac_f64_dpp %0, %1, %" #op " row_newbcast:" #k " row_mask:0xf bank_mask:0xf\n"
  asm("s_nop 4\n" DMPC_FMAC16(2, 0) DMPC_FMAC16(3, 1) DMPC_FMAC16(4, 2) DMPC_FMAC16(5, 3) DMPC_FMAC16(6, 4)
          DMPC_FMAC16(7, 5) DMPC_FMAC16(8, 6) DMPC_FMAC16(9, 7) DMPC_FMAC16(10, 8) DMPC_FMAC16(11, 9)
              DMPC_FMAC16(12, 10) DMPC_FMAC16(13, 11) DMPC_FMAC16(14, 12) DMPC_FMAC16(15, 13) DMPC_FMAC16(16, 14)
                  DMPC_FMAC16(17, 15)
      : "+v"(acc)
      : "v"(b), "v"(s[0]), "v"(s[1]), "v"(s[2]), "v"(s[3]), "v"(s[4]), "v"(s[5]), "v"(s[6]), "v"(s[7]),
        "v"(s[8]), "v"(s[9]), "v"(s[10]), "v"(s[11]), "v"(s[12]), "v"(s[13]), "v"(s[14]), "v"(s[15]));
#undef DMPC_FMAC16
  return acc;
}
template <int M, int j = 0, class AT>
__device__ __forceinline__ double gdot64_rows(const AT* a, const double (&b)[4], int salt, double acc) {
  if constexpr (j == M) {
    return acc;
  } else if constexpr (j % 16 == 0 && j + 16 <= M) {
    double s[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) s[k] = cvt_in_loop(a[j + k], salt);
    acc = fmac_row16(acc, b[j / 16], s);
    return gdot64_rows<M, j + 16>(a, b, salt, acc);
  } else {
    dpp_fmac16<j % 16>(acc, b[j / 16], cvt_in_loop(a[j], salt));
    return gdot64_rows<M, j + 1>(a, b, salt, acc);
  }
}

// sum_j a[j] * v_j where v_j lives in lane j of the group (j < M <= NX).
template <int G, int M, int j = 0, class AT>
__device__ __forceinline__ double gdot_chain(const AT* a, double v, int lane, int salt, double acc) {
  if constexpr (j == M) {
    return acc;
  } else {
    acc += cvt_in_loop(a[j], salt) * gbcast<G, j>(v, lane);
    // keep the scheduler from converting and broadcasting a whole row ahead
    // of its FMAs (each in-flight term holds four VGPRs)
    if constexpr (j % 4 == 3) __builtin_amdgcn_sched_barrier(0);
    return gdot_chain<G, M, j + 1>(a, v, lane, salt, acc);
  }
}

// acc + sum_{j<13} a[j] * v_j for 16-lane groups with a[] in float64
// registers: each term is one v_fmac_f64 whose first operand is read through
// DPP row_newbcast:j (lane j of the 16-lane row), so the broadcast costs no
// instruction.  One asm block keeps the chain free of compiler-inserted
// writes to v; the leading s_nop 4 covers the DPP read-after-VALU-write wait
// states (2 for a VGPR, 5 after a VALU write of EXEC).  Terms are summed in
// order j = 0..12 with one rounding each, as the generic chain does.
__device__ __forceinline__ double gdot16x13(const double* a, double v, double acc) {
#define DMPC_FMAC_BCAST(op, j) "v_fmac_f64_dpp %0, %1, %" #op " row_newbcast:" #j " row_mask:0xf bank_mask:0xf\n"
  asm("s_nop 4\n" DMPC_FMAC_BCAST(2, 0) DMPC_FMAC_BCAST(3, 1) DMPC_FMAC_BCAST(4, 2) DMPC_FMAC_BCAST(5, 3)
          DMPC_FMAC_BCAST(6, 4) DMPC_FMAC_BCAST(7, 5) DMPC_FMAC_BCAST(8, 6) DMPC_FMAC_BCAST(9, 7)
              DMPC_FMAC_BCAST(10, 8) DMPC_FMAC_BCAST(11, 9) DMPC_FMAC_BCAST(12, 10) DMPC_FMAC_BCAST(13, 11)
                  DMPC_FMAC_BCAST(14, 12)
      : "+v"(acc)
      : "v"(v), "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(a[5]), "v"(a[6]), "v"(a[7]),
        "v"(a[8]), "v"(a[9]), "v"(a[10]), "v"(a[11]), "v"(a[12]));
#undef DMPC_FMAC_BCAST
  return acc;
}

// Two independent gdot16x13 chains in one instruction stream, terms of the
// two chains alternating: acc0 += sum_j a0[j] v0_j, acc1 += sum_j a1[j] v1_j,
// each chain summed in order j = 0..12 with one rounding per term (the same
// arithmetic as two gdot16x13 calls).  A lane with two time blocks, or two
// products of one block that do not depend on each other, then keeps two
// FMAs in flight, which the dependent single chain cannot.  Terms j < 7 and
// j >= 7 are two asm blocks so that the caller can load the second half of
// its operands after the first half has been consumed.
#define DMPC_FMAC_PAIR(acc, vv, op, j) "v_fmac_f64_dpp %" #acc ", %" #vv ", %" #op " row_newbcast:" #j " row_mask:0xf bank_mask:0xf\n"
__device__ __forceinline__ void gdot16x13_pair_lo(const double* a0, double v0, const double* a1, double v1,
                                                  double& acc0, double& acc1) {
  asm("s_nop 4\n" DMPC_FMAC_PAIR(0, 2, 4, 0) DMPC_FMAC_PAIR(1, 3, 11, 0) DMPC_FMAC_PAIR(0, 2, 5, 1)
          DMPC_FMAC_PAIR(1, 3, 12, 1) DMPC_FMAC_PAIR(0, 2, 6, 2) DMPC_FMAC_PAIR(1, 3, 13, 2)
              DMPC_FMAC_PAIR(0, 2, 7, 3) DMPC_FMAC_PAIR(1, 3, 14, 3) DMPC_FMAC_PAIR(0, 2, 8, 4)
                  DMPC_FMAC_PAIR(1, 3, 15, 4) DMPC_FMAC_PAIR(0, 2, 9, 5) DMPC_FMAC_PAIR(1, 3, 16, 5)
                      DMPC_FMAC_PAIR(0, 2, 10, 6) DMPC_FMAC_PAIR(1, 3, 17, 6)
      : "+&v"(acc0), "+&v"(acc1)
      : "v"(v0), "v"(v1), "v"(a0[0]), "v"(a0[1]), "v"(a0[2]), "v"(a0[3]), "v"(a0[4]), "v"(a0[5]), "v"(a0[6]),
        "v"(a1[0]), "v"(a1[1]), "v"(a1[2]), "v"(a1[3]), "v"(a1[4]), "v"(a1[5]), "v"(a1[6]));
}
// terms j = 7..12; a0, a1 point at element 7
__device__ __forceinline__ void gdot16x13_pair_hi(const double* a0, double v0, const double* a1, double v1,
                                                  double& acc0, double& acc1) {
  asm("s_nop 4\n" DMPC_FMAC_PAIR(0, 2, 4, 7) DMPC_FMAC_PAIR(1, 3, 10, 7) DMPC_FMAC_PAIR(0, 2, 5, 8)
          DMPC_FMAC_PAIR(1, 3, 11, 8) DMPC_FMAC_PAIR(0, 2, 6, 9) DMPC_FMAC_PAIR(1, 3, 12, 9)
              DMPC_FMAC_PAIR(0, 2, 7, 10) DMPC_FMAC_PAIR(1, 3, 13, 10) DMPC_FMAC_PAIR(0, 2, 8, 11)
                  DMPC_FMAC_PAIR(1, 3, 14, 11) DMPC_FMAC_PAIR(0, 2, 9, 12) DMPC_FMAC_PAIR(1, 3, 15, 12)
      : "+&v"(acc0), "+&v"(acc1)
      : "v"(v0), "v"(v1), "v"(a0[0]), "v"(a0[1]), "v"(a0[2]), "v"(a0[3]), "v"(a0[4]), "v"(a0[5]), "v"(a1[0]),
        "v"(a1[1]), "v"(a1[2]), "v"(a1[3]), "v"(a1[4]), "v"(a1[5]));
}
#undef DMPC_FMAC_PAIR
__device__ __forceinline__ void gdot16x13_pair(const double* a0, double v0, const double* a1, double v1,
                                               double& acc0, double& acc1) {
  gdot16x13_pair_lo(a0, v0, a1, v1, acc0, acc1);
  gdot16x13_pair_hi(a0 + 7, v0, a1 + 7, v1, acc0, acc1);
}

// As gdot16x13 for float32 a[]: each term converts a[j] into a scratch pair
// and feeds it to the DPP FMA (the conversion writes the non-DPP operand, so
// it needs no wait state).  Same terms, order and roundings as the generic
// chain (exact conversion, one FMA rounding per term).
__device__ __forceinline__ double gdot16x13f(const float* a, double v, double acc) {
  double tmp;
#define DMPC_CVT_FMAC(op, j)                                              \
  "v_cvt_f64_f32 %1, %" #op "\n"                                          \
  "v_fmac_f64_dpp %0, %2, %1 row_newbcast:" #j " row_mask:0xf bank_mask:0xf\n"
  asm("s_nop 4\n" DMPC_CVT_FMAC(3, 0) DMPC_CVT_FMAC(4, 1) DMPC_CVT_FMAC(5, 2) DMPC_CVT_FMAC(6, 3)
          DMPC_CVT_FMAC(7, 4) DMPC_CVT_FMAC(8, 5) DMPC_CVT_FMAC(9, 6) DMPC_CVT_FMAC(10, 7)
              DMPC_CVT_FMAC(11, 8) DMPC_CVT_FMAC(12, 9) DMPC_CVT_FMAC(13, 10) DMPC_CVT_FMAC(14, 11)
                  DMPC_CVT_FMAC(15, 12)
      : "+v"(acc), "=&v"(tmp)
      : "v"(v), "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(a[5]), "v"(a[6]), "v"(a[7]),
        "v"(a[8]), "v"(a[9]), "v"(a[10]), "v"(a[11]), "v"(a[12]));
#undef DMPC_CVT_FMAC
  return acc;
}

// acc + sum_{j<M} a[j] * v_j for 8-lane groups, which sit two per 16-lane
// row.  float64 DPP takes only row_newbcast (lane j of the row to all 16
// lanes) and ignores bank masks, so the chain runs twice with EXEC narrowed
// (ANDed with the caller's mask and restored in the same asm block): lanes
// 0-7 of every row with row_newbcast:j, then lanes 8-15 with
// row_newbcast:8+j.  One v_fmac_f64_dpp per term and lane half (two
// instructions against five for gbcast's four 32-bit moves and an FMA); the
// s_nop 4 after each EXEC write covers the SALU-EXEC-to-DPP wait states and
// the first also the DPP source's read-after-write; SCC is clobbered.  float
// a[] converts each term into a scratch pair first, as gdot16x13f does.  Same
// terms, order and roundings as gdot_chain.
constexpr unsigned long long kLowHalves = 0x00ff00ff00ff00ffull;  // lanes 0-7 of each 16-lane row

template <int M>
__device__ __forceinline__ double gdot8d_fused(const double* a, double v, double acc) {
  unsigned long long save;
  if constexpr (M == 1) {
    asm volatile("s_mov_b64 %1, exec\n"
        "s_and_b64 exec, exec, %2\n"
        "s_nop 4\n"
        "v_fmac_f64_dpp %0, %4, %3 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "s_andn2_b64 exec, %1, %2\n"
        "s_nop 4\n"
        "v_fmac_f64_dpp %0, %4, %3 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "s_mov_b64 exec, %1\n"
        "s_nop 4\n"
        : "+v"(acc), "=&s"(save)
        : "s"(kLowHalves), "v"(a[0]), "v"(v) : "scc");
  } else if constexpr (M == 2) {
    asm volatile("s_mov_b64 %1, exec\n"
        "s_and_b64 exec, exec, %2\n"
        "s_nop 4\n"
        "v_fmac_f64_dpp %0, %5, %3 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %5, %4 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "s_andn2_b64 exec, %1, %2\n"
        "s_nop 4\n"
        "v_fmac_f64_dpp %0, %5, %3 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %5, %4 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "s_mov_b64 exec, %1\n"
        "s_nop 4\n"
        : "+v"(acc), "=&s"(save)
        : "s"(kLowHalves), "v"(a[0]), "v"(a[1]), "v"(v) : "scc");
  } else if constexpr (M == 3) {
    asm volatile("s_mov_b64 %1, exec\n"
        "s_and_b64 exec, exec, %2\n"
        "s_nop 4\n"
        "v_fmac_f64_dpp %0, %6, %3 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %6, %4 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %6, %5 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "s_andn2_b64 exec, %1, %2\n"
        "s_nop 4\n"
        "v_fmac_f64_dpp %0, %6, %3 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %6, %4 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %6, %5 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "s_mov_b64 exec, %1\n"
        "s_nop 4\n"
        : "+v"(acc), "=&s"(save)
        : "s"(kLowHalves), "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(v) : "scc");
  } else if constexpr (M == 4) {
    asm volatile("s_mov_b64 %1, exec\n"
        "s_and_b64 exec, exec, %2\n"
        "s_nop 4\n"
        "v_fmac_f64_dpp %0, %7, %3 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %7, %4 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %7, %5 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %7, %6 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "s_andn2_b64 exec, %1, %2\n"
        "s_nop 4\n"
        "v_fmac_f64_dpp %0, %7, %3 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %7, %4 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %7, %5 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %7, %6 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "s_mov_b64 exec, %1\n"
        "s_nop 4\n"
        : "+v"(acc), "=&s"(save)
        : "s"(kLowHalves), "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(v) : "scc");
  } else if constexpr (M == 5) {
    asm volatile("s_mov_b64 %1, exec\n"
        "s_and_b64 exec, exec, %2\n"
        "s_nop 4\n"
        "v_fmac_f64_dpp %0, %8, %3 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %8, %4 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %8, %5 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %8, %6 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %8, %7 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "s_andn2_b64 exec, %1, %2\n"
        "s_nop 4\n"
        "v_fmac_f64_dpp %0, %8, %3 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %8, %4 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %8, %5 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %8, %6 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %8, %7 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "s_mov_b64 exec, %1\n"
        "s_nop 4\n"
        : "+v"(acc), "=&s"(save)
        : "s"(kLowHalves), "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(v) : "scc");
  } else if constexpr (M == 6) {
    asm volatile("s_mov_b64 %1, exec\n"
        "s_and_b64 exec, exec, %2\n"
        "s_nop 4\n"
        "v_fmac_f64_dpp %0, %9, %3 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %9, %4 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %9, %5 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %9, %6 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %9, %7 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %9, %8 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "s_andn2_b64 exec, %1, %2\n"
        "s_nop 4\n"
        "v_fmac_f64_dpp %0, %9, %3 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %9, %4 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %9, %5 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %9, %6 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %9, %7 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %9, %8 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "s_mov_b64 exec, %1\n"
        "s_nop 4\n"
        : "+v"(acc), "=&s"(save)
        : "s"(kLowHalves), "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(a[5]), "v"(v) : "scc");
  } else if constexpr (M == 7) {
    asm volatile("s_mov_b64 %1, exec\n"
        "s_and_b64 exec, exec, %2\n"
        "s_nop 4\n"
        "v_fmac_f64_dpp %0, %10, %3 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %10, %4 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %10, %5 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %10, %6 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %10, %7 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %10, %8 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %10, %9 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "s_andn2_b64 exec, %1, %2\n"
        "s_nop 4\n"
        "v_fmac_f64_dpp %0, %10, %3 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %10, %4 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %10, %5 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %10, %6 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %10, %7 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %10, %8 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %10, %9 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "s_mov_b64 exec, %1\n"
        "s_nop 4\n"
        : "+v"(acc), "=&s"(save)
        : "s"(kLowHalves), "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(a[5]), "v"(a[6]), "v"(v) : "scc");
  } else if constexpr (M == 8) {
    asm volatile("s_mov_b64 %1, exec\n"
        "s_and_b64 exec, exec, %2\n"
        "s_nop 4\n"
        "v_fmac_f64_dpp %0, %11, %3 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %11, %4 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %11, %5 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %11, %6 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %11, %7 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %11, %8 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %11, %9 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %11, %10 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "s_andn2_b64 exec, %1, %2\n"
        "s_nop 4\n"
        "v_fmac_f64_dpp %0, %11, %3 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %11, %4 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %11, %5 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %11, %6 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %11, %7 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %11, %8 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %11, %9 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %11, %10 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        "s_mov_b64 exec, %1\n"
        "s_nop 4\n"
        : "+v"(acc), "=&s"(save)
        : "s"(kLowHalves), "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(a[5]), "v"(a[6]), "v"(a[7]), "v"(v) : "scc");
  }
  return acc;
}

template <int M>
__device__ __forceinline__ double gdot8f_fused(const float* a, double v, double acc) {
  unsigned long long save;
  double tmp;
  if constexpr (M == 1) {
    asm volatile("s_mov_b64 %1, exec\n"
        "s_and_b64 exec, exec, %3\n"
        "s_nop 4\n"
        "v_cvt_f64_f32 %2, %4\n"
        "v_fmac_f64_dpp %0, %5, %2 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "s_andn2_b64 exec, %1, %3\n"
        "s_nop 4\n"
        "v_cvt_f64_f32 %2, %4\n"
        "v_fmac_f64_dpp %0, %5, %2 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "s_mov_b64 exec, %1\n"
        "s_nop 4\n"
        : "+v"(acc), "=&s"(save), "=&v"(tmp)
        : "s"(kLowHalves), "v"(a[0]), "v"(v) : "scc");
  } else if constexpr (M == 2) {
    asm volatile("s_mov_b64 %1, exec\n"
        "s_and_b64 exec, exec, %3\n"
        "s_nop 4\n"
        "v_cvt_f64_f32 %2, %4\n"
        "v_fmac_f64_dpp %0, %6, %2 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %5\n"
        "v_fmac_f64_dpp %0, %6, %2 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "s_andn2_b64 exec, %1, %3\n"
        "s_nop 4\n"
        "v_cvt_f64_f32 %2, %4\n"
        "v_fmac_f64_dpp %0, %6, %2 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %5\n"
        "v_fmac_f64_dpp %0, %6, %2 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "s_mov_b64 exec, %1\n"
        "s_nop 4\n"
        : "+v"(acc), "=&s"(save), "=&v"(tmp)
        : "s"(kLowHalves), "v"(a[0]), "v"(a[1]), "v"(v) : "scc");
  } else if constexpr (M == 3) {
    asm volatile("s_mov_b64 %1, exec\n"
        "s_and_b64 exec, exec, %3\n"
        "s_nop 4\n"
        "v_cvt_f64_f32 %2, %4\n"
        "v_fmac_f64_dpp %0, %7, %2 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %5\n"
        "v_fmac_f64_dpp %0, %7, %2 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %6\n"
        "v_fmac_f64_dpp %0, %7, %2 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "s_andn2_b64 exec, %1, %3\n"
        "s_nop 4\n"
        "v_cvt_f64_f32 %2, %4\n"
        "v_fmac_f64_dpp %0, %7, %2 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %5\n"
        "v_fmac_f64_dpp %0, %7, %2 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %6\n"
        "v_fmac_f64_dpp %0, %7, %2 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "s_mov_b64 exec, %1\n"
        "s_nop 4\n"
        : "+v"(acc), "=&s"(save), "=&v"(tmp)
        : "s"(kLowHalves), "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(v) : "scc");
  } else if constexpr (M == 4) {
    asm volatile("s_mov_b64 %1, exec\n"
        "s_and_b64 exec, exec, %3\n"
        "s_nop 4\n"
        "v_cvt_f64_f32 %2, %4\n"
        "v_fmac_f64_dpp %0, %8, %2 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %5\n"
        "v_fmac_f64_dpp %0, %8, %2 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %6\n"
        "v_fmac_f64_dpp %0, %8, %2 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %7\n"
        "v_fmac_f64_dpp %0, %8, %2 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "s_andn2_b64 exec, %1, %3\n"
        "s_nop 4\n"
        "v_cvt_f64_f32 %2, %4\n"
        "v_fmac_f64_dpp %0, %8, %2 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %5\n"
        "v_fmac_f64_dpp %0, %8, %2 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %6\n"
        "v_fmac_f64_dpp %0, %8, %2 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %7\n"
        "v_fmac_f64_dpp %0, %8, %2 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "s_mov_b64 exec, %1\n"
        "s_nop 4\n"
        : "+v"(acc), "=&s"(save), "=&v"(tmp)
        : "s"(kLowHalves), "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(v) : "scc");
  } else if constexpr (M == 5) {
    asm volatile("s_mov_b64 %1, exec\n"
        "s_and_b64 exec, exec, %3\n"
        "s_nop 4\n"
        "v_cvt_f64_f32 %2, %4\n"
        "v_fmac_f64_dpp %0, %9, %2 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %5\n"
        "v_fmac_f64_dpp %0, %9, %2 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %6\n"
        "v_fmac_f64_dpp %0, %9, %2 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %7\n"
        "v_fmac_f64_dpp %0, %9, %2 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %8\n"
        "v_fmac_f64_dpp %0, %9, %2 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "s_andn2_b64 exec, %1, %3\n"
        "s_nop 4\n"
        "v_cvt_f64_f32 %2, %4\n"
        "v_fmac_f64_dpp %0, %9, %2 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %5\n"
        "v_fmac_f64_dpp %0, %9, %2 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %6\n"
        "v_fmac_f64_dpp %0, %9, %2 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %7\n"
        "v_fmac_f64_dpp %0, %9, %2 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %8\n"
        "v_fmac_f64_dpp %0, %9, %2 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "s_mov_b64 exec, %1\n"
        "s_nop 4\n"
        : "+v"(acc), "=&s"(save), "=&v"(tmp)
        : "s"(kLowHalves), "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(v) : "scc");
  } else if constexpr (M == 6) {
    asm volatile("s_mov_b64 %1, exec\n"
        "s_and_b64 exec, exec, %3\n"
        "s_nop 4\n"
        "v_cvt_f64_f32 %2, %4\n"
        "v_fmac_f64_dpp %0, %10, %2 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %5\n"
        "v_fmac_f64_dpp %0, %10, %2 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %6\n"
        "v_fmac_f64_dpp %0, %10, %2 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %7\n"
        "v_fmac_f64_dpp %0, %10, %2 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %8\n"
        "v_fmac_f64_dpp %0, %10, %2 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %9\n"
        "v_fmac_f64_dpp %0, %10, %2 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "s_andn2_b64 exec, %1, %3\n"
        "s_nop 4\n"
        "v_cvt_f64_f32 %2, %4\n"
        "v_fmac_f64_dpp %0, %10, %2 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %5\n"
        "v_fmac_f64_dpp %0, %10, %2 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %6\n"
        "v_fmac_f64_dpp %0, %10, %2 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %7\n"
        "v_fmac_f64_dpp %0, %10, %2 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %8\n"
        "v_fmac_f64_dpp %0, %10, %2 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %9\n"
        "v_fmac_f64_dpp %0, %10, %2 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "s_mov_b64 exec, %1\n"
        "s_nop 4\n"
        : "+v"(acc), "=&s"(save), "=&v"(tmp)
        : "s"(kLowHalves), "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(a[5]), "v"(v) : "scc");
  } else if constexpr (M == 7) {
    asm volatile("s_mov_b64 %1, exec\n"
        "s_and_b64 exec, exec, %3\n"
        "s_nop 4\n"
        "v_cvt_f64_f32 %2, %4\n"
        "v_fmac_f64_dpp %0, %11, %2 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %5\n"
        "v_fmac_f64_dpp %0, %11, %2 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %6\n"
        "v_fmac_f64_dpp %0, %11, %2 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %7\n"
        "v_fmac_f64_dpp %0, %11, %2 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %8\n"
        "v_fmac_f64_dpp %0, %11, %2 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %9\n"
        "v_fmac_f64_dpp %0, %11, %2 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %10\n"
        "v_fmac_f64_dpp %0, %11, %2 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "s_andn2_b64 exec, %1, %3\n"
        "s_nop 4\n"
        "v_cvt_f64_f32 %2, %4\n"
        "v_fmac_f64_dpp %0, %11, %2 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %5\n"
        "v_fmac_f64_dpp %0, %11, %2 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %6\n"
        "v_fmac_f64_dpp %0, %11, %2 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %7\n"
        "v_fmac_f64_dpp %0, %11, %2 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %8\n"
        "v_fmac_f64_dpp %0, %11, %2 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %9\n"
        "v_fmac_f64_dpp %0, %11, %2 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %10\n"
        "v_fmac_f64_dpp %0, %11, %2 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "s_mov_b64 exec, %1\n"
        "s_nop 4\n"
        : "+v"(acc), "=&s"(save), "=&v"(tmp)
        : "s"(kLowHalves), "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(a[5]), "v"(a[6]), "v"(v) : "scc");
  } else if constexpr (M == 8) {
    asm volatile("s_mov_b64 %1, exec\n"
        "s_and_b64 exec, exec, %3\n"
        "s_nop 4\n"
        "v_cvt_f64_f32 %2, %4\n"
        "v_fmac_f64_dpp %0, %12, %2 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %5\n"
        "v_fmac_f64_dpp %0, %12, %2 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %6\n"
        "v_fmac_f64_dpp %0, %12, %2 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %7\n"
        "v_fmac_f64_dpp %0, %12, %2 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %8\n"
        "v_fmac_f64_dpp %0, %12, %2 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %9\n"
        "v_fmac_f64_dpp %0, %12, %2 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %10\n"
        "v_fmac_f64_dpp %0, %12, %2 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %11\n"
        "v_fmac_f64_dpp %0, %12, %2 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "s_andn2_b64 exec, %1, %3\n"
        "s_nop 4\n"
        "v_cvt_f64_f32 %2, %4\n"
        "v_fmac_f64_dpp %0, %12, %2 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %5\n"
        "v_fmac_f64_dpp %0, %12, %2 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %6\n"
        "v_fmac_f64_dpp %0, %12, %2 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %7\n"
        "v_fmac_f64_dpp %0, %12, %2 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %8\n"
        "v_fmac_f64_dpp %0, %12, %2 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %9\n"
        "v_fmac_f64_dpp %0, %12, %2 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %10\n"
        "v_fmac_f64_dpp %0, %12, %2 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_cvt_f64_f32 %2, %11\n"
        "v_fmac_f64_dpp %0, %12, %2 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        "s_mov_b64 exec, %1\n"
        "s_nop 4\n"
        : "+v"(acc), "=&s"(save), "=&v"(tmp)
        : "s"(kLowHalves), "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(a[5]), "v"(a[6]), "v"(a[7]), "v"(v) : "scc");
  }
  return acc;
}

// ---------------------------------------------------------------------------
// B-term of a 16-lane group with four columns, by reduce-scatter and
// all-gather instead of four group all-reduces.  Lane i of the group belongs
// to quad q = (i >> 2) & 3 and keeps row i of B with its columns permuted:
// slot r holds B[i][q ^ r].  Then
//   rs16:     every lane of quad q gets s_q = sum_j B[j][q] v_j
//             (partners i^8, i^7, i^1, i^2 via row_ror:8, row_half_mirror
//             and two quad permutations; each step halves what is sent);
//   gather16: sum_k B[i][k] y_k with y_k held by quad k (quads q^1, q^2, q^3
//             are reached with row_half_mirror, row_ror:8, row_mirror).
// 19 + 10 VALU instructions against 64 for four DPP all-reduces.
constexpr int kDppRor8 = 0x128, kDppHalfMirror = 0x141, kDppMirror = 0x140;

__device__ __forceinline__ int quad16(int lane) { return (lane >> 2) & 3; }

template <class BT>
__device__ __forceinline__ double rs16(const BT (&bb)[4], double v, int salt) {
  double t0 = cvt_in_loop(bb[0], salt) * v, t1 = cvt_in_loop(bb[1], salt) * v;
  const double s2 = cvt_in_loop(bb[2], salt) * v, s3 = cvt_in_loop(bb[3], salt) * v;
  t0 += dpp_mov<kDppRor8>(s2);
  t1 += dpp_mov<kDppRor8>(s3);
  t0 += dpp_mov<kDppHalfMirror>(t1);
  t0 += dpp_mov<0xB1>(t0);  // quad_perm [1,0,3,2]
  t0 += dpp_mov<0x4E>(t0);  // quad_perm [2,3,0,1]
  return t0;
}

template <class BT>
__device__ __forceinline__ double gather16(const BT (&bb)[4], double y, int salt) {
  double g = cvt_in_loop(bb[0], salt) * y;
  g += cvt_in_loop(bb[1], salt) * dpp_mov<kDppHalfMirror>(y);
  g += cvt_in_loop(bb[2], salt) * dpp_mov<kDppRor8>(y);
  g += cvt_in_loop(bb[3], salt) * dpp_mov<kDppMirror>(y);
  return g;
}

template <int G, int M, class AT>
__device__ __forceinline__ double gdot(const AT* a, double v, int lane, int salt) {
  if constexpr (G == 16 && M == 13 && std::is_same<AT, double>::value) {
    return gdot16x13(a, v, 0.0);
  } else if constexpr (G == 16 && M == 13 && std::is_same<AT, float>::value) {
    return gdot16x13f(a, v, 0.0);
  } else if constexpr (G == 8 && M <= 8 && std::is_same<AT, double>::value) {
    return gdot8d_fused<M>(a, v, 0.0);
  } else if constexpr (G == 8 && M <= 8 && std::is_same<AT, float>::value) {
    return gdot8f_fused<M>(a, v, 0.0);
  } else if constexpr (G == 64) {
    double b[4];
    rows_to_all(v, b);
    return gdot64_rows<M>(a, b, salt, 0.0);
  } else {
    return gdot_chain<G, M>(a, v, lane, salt, 0.0);
  }
}
template <int G, int M, int j = 0>
__device__ __forceinline__ double gdotd(const double* a, double v, int lane, double acc = 0.0) {
  if constexpr (j == 0 && G == 8 && M <= 8) {
    return gdot8d_fused<M>(a, v, acc);
  } else if constexpr (j == M) {
    return acc;
  } else {
    acc += a[j] * gbcast<G, j>(v, lane);
    if constexpr (j % 4 == 3) __builtin_amdgcn_sched_barrier(0);
    return gdotd<G, M, j + 1>(a, v, lane, acc);
  }
}


// sum_j a[j * STRIDE] * v_j (a in LDS), v_j in lane j of the group.
template <int G, int M, int STRIDE, int j = 0>
__device__ __forceinline__ double gdot_strided(const double* a, double v, int lane, double acc = 0.0) {
  if constexpr (j == M) {
    return acc;
  } else {
    acc += a[j * STRIDE] * gbcast<G, j>(v, lane);
    if constexpr (j % 4 == 3) __builtin_amdgcn_sched_barrier(0);
    return gdot_strided<G, M, STRIDE, j + 1>(a, v, lane, acc);
  }
}

}  // namespace
}  // namespace dmpc
