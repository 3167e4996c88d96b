// Sums over the 64 lanes of a wave of several values per lane, in registers.
//
// wave_reduce_scatter<N>(x): every lane holds x[0..N-1] (N <= 64); after six
// exchange steps lane l holds the sum over all 64 lanes of one value,
// x[wave_rs_index(N, l)] (an index of -1 marks a lane left with padding).
// Each step pairs every lane with a partner in the other half of a lane
// group and halves the values each lane holds: of a pair (x[j], x[j + H]),
// the lane in the lower half of the group keeps x[j] plus its partner's
// x[j], the lane in the upper half keeps x[j + H] plus its partner's
// x[j + H] (a list of odd length is padded with one zero at its end).  The
// groups, from the widest:
//   lane bit 5: v_permlane32_swap (the two 32-lane halves),
//   lane bit 4: v_permlane16_swap (16-lane rows 0/1 and 2/3),
//   lane bit 3: DPP row_mirror (lane i of a row with lane 15 - i),
//   lane bit 2: DPP row_half_mirror (lane i of an 8-lane half with 7 - i),
//   lane bit 1: DPP quad_perm [2, 3, 0, 1],
//   lane bit 0: DPP quad_perm [1, 0, 3, 2].
// N values cost about N/2 + N/4 + ... pair steps of one to three VALU
// operations each and no LDS traffic; the tree is fixed, so every run adds
// the same terms in the same order.  All 64 lanes must be active.
#pragma once

#include <hip/hip_runtime.h>

namespace kh {

// The value that lane `lane` holds after wave_reduce_scatter<n>, or -1.
__host__ __device__ constexpr int wave_rs_index(int n, int lane) {
  int start = 0, cnt = n, m = n;
  for (int s = 5; s >= 0; --s) {
    const int h = (m + 1) / 2;  // pairs after padding m to even
    if ((lane >> s) & 1) {
      start += h;
      cnt = cnt > h ? cnt - h : 0;
    } else {
      cnt = cnt < h ? cnt : h;
    }
    m = h;
  }
  return cnt > 0 ? start : -1;
}

__device__ __forceinline__ unsigned lo32(double x) {
  return (unsigned)__double_as_longlong(x);
}
__device__ __forceinline__ unsigned hi32(double x) {
  return (unsigned)((unsigned long long)__double_as_longlong(x) >> 32);
}
__device__ __forceinline__ double join64(unsigned lo, unsigned hi) {
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// One pair (a, b) of step S (lane bit S): see the header comment.
template <int S>
__device__ __forceinline__ double rs_pair(double a, double b, int lane) {
  if constexpr (S == 5 || S == 4) {
    // permlaneN_swap(p, q) exchanges the upper group's p with the lower
    // group's q, so afterwards p + q is the a-sum in the lower group and
    // the b-sum in the upper one
    auto sl = S == 5 ? __builtin_amdgcn_permlane32_swap(lo32(a), lo32(b), false, false)
                     : __builtin_amdgcn_permlane16_swap(lo32(a), lo32(b), false, false);
    auto sh = S == 5 ? __builtin_amdgcn_permlane32_swap(hi32(a), hi32(b), false, false)
                     : __builtin_amdgcn_permlane16_swap(hi32(a), hi32(b), false, false);
    return join64(sl[0], sh[0]) + join64(sl[1], sh[1]);
  } else {
    constexpr int CTRL = S == 3 ? 0x140 : S == 2 ? 0x141 : S == 1 ? 0x4e : 0xb1;
    const bool up = (lane >> S) & 1;
    const double send = up ? a : b;  // what the partner keeps
    const double keep = up ? b : a;
    const double got = join64(
        (unsigned)__builtin_amdgcn_update_dpp(0, (int)lo32(send), CTRL, 0xf, 0xf, false),
        (unsigned)__builtin_amdgcn_update_dpp(0, (int)hi32(send), CTRL, 0xf, 0xf, false));
    return keep + got;
  }
}

template <int S, int M>
struct WaveRs {
  static constexpr int H = (M + 1) / 2;
  __device__ static __forceinline__ double run(const double (&x)[M], int lane) {
    double y[H];
#pragma unroll
    for (int j = 0; j < H; ++j) y[j] = rs_pair<S>(x[j], j + H < M ? x[j + H] : 0.0, lane);
    if constexpr (S == 0) {
      static_assert(H == 1, "at most 64 values");
      return y[0];
    } else {
      return WaveRs<S - 1, H>::run(y, lane);
    }
  }
};

// The 32-bit value x of this lane's partner at step S of an all-reduce
// over the wave: lane ^ 1, lane ^ 2 (quad permutations), the mirror lane of
// an 8-lane half row, of a 16-lane row, then lane ^ 16 and lane ^ 32 (the
// permlane swaps).  After the six steps, combining each lane's value with
// its partner's at every step, every lane holds the combination of all 64.
template <int S>
__device__ __forceinline__ unsigned wave_partner_b32(unsigned x) {
  if constexpr (S == 0) return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0xb1, 0xf, 0xf, false);
  else if constexpr (S == 1) return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x4e, 0xf, 0xf, false);
  else if constexpr (S == 2) return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xf, 0xf, false);
  else if constexpr (S == 3) return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x140, 0xf, 0xf, false);
  else {
    // permlaneN_swap(x, x): the lower group's copy of the upper group's x
    // in the second result, the upper group's copy of the lower's in the first
    auto r = S == 4 ? __builtin_amdgcn_permlane16_swap(x, x, false, false)
                    : __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return ((threadIdx.x >> S) & 1) ? r[0] : r[1];
  }
}

template <int N>
__device__ __forceinline__ double wave_reduce_scatter(const double (&x)[N]) {
  static_assert(N >= 1 && N <= 64, "1 to 64 values per lane");
  return WaveRs<5, N>::run(x, (int)(threadIdx.x & 63));
}

}  // namespace kh
