// Wave64 cross-lane helpers for gfx950 (CDNA4): prefix sums, counts and reductions built from
// DPP row shifts, v_readlane and ballot/mbcnt, so they use no LDS permute round trip. Every lane of
// the wave must be active when these are called.
#pragma once
#include <hip/hip_runtime.h>

namespace lagpm_wave {

// Cross-lane steps below use DPP (data-parallel primitives: a VALU operand read from another lane
// of the same 16-lane row) and v_readlane, never the LDS permute path, so a walk round's prefix
// sums and reductions wait on no LDS round trip. DPP controls: row_shr:n = 0x110 + n.
template <int CTRL>
__device__ __forceinline__ int dpp_shr_zero(int x) {  // lane i of a row gets lane i - n of the row, 0 past its start
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xf, 0xf, true);
}
template <int CTRL>
__device__ __forceinline__ int dpp_shr_keep(int x) {  // lane i of a row gets lane i - n of the row, x past its start
    return __builtin_amdgcn_update_dpp(x, x, CTRL, 0xf, 0xf, false);
}
// Exclusive wave prefix sum. The total is read from lane 63 into a scalar register, so every
// loop or branch the walk takes on it is wave-uniform by construction (scalar branches): the
// compiler can then never split the wave's lanes across different iterations of the walk.
// Inclusive scan within each row (row_shr 1, 2, 4, 8), then the row totals from lanes 15, 31
// and 47 (scalars) are added to the later rows.
__device__ __forceinline__ int wave_excl_scan(int v, int lane, int& total) {
    int x = v;
    x += dpp_shr_zero<0x111>(x);
    x += dpp_shr_zero<0x112>(x);
    x += dpp_shr_zero<0x114>(x);
    x += dpp_shr_zero<0x118>(x);
    const int r0 = __builtin_amdgcn_readlane(x, 15);
    const int r1 = r0 + __builtin_amdgcn_readlane(x, 31);
    const int r2 = r1 + __builtin_amdgcn_readlane(x, 47);
    const int row = lane >> 4;
    x += row == 0 ? 0 : (row == 1 ? r0 : (row == 2 ? r1 : r2));
    total = __builtin_amdgcn_readlane(x, 63);
    return x - v;
}
// Exclusive count of the lanes below this one whose flag is set (ballot + mbcnt), and the total.
__device__ __forceinline__ int wave_excl_count(bool f, int& total) {
    const unsigned long long m = __ballot(f);
    total = __builtin_amdgcn_readfirstlane(__popcll(m));
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}
// Wave minimum / maximum: after row_shr 1, 2, 4, 8 lane 15 of each row holds the row's extreme;
// the four row results are combined from scalars.
__device__ __forceinline__ float wave_minf(float v) {
    int x = __float_as_int(v);
    x = __float_as_int(fminf(__int_as_float(x), __int_as_float(dpp_shr_keep<0x111>(x))));
    x = __float_as_int(fminf(__int_as_float(x), __int_as_float(dpp_shr_keep<0x112>(x))));
    x = __float_as_int(fminf(__int_as_float(x), __int_as_float(dpp_shr_keep<0x114>(x))));
    x = __float_as_int(fminf(__int_as_float(x), __int_as_float(dpp_shr_keep<0x118>(x))));
    const float a = __int_as_float(__builtin_amdgcn_readlane(x, 15)), b = __int_as_float(__builtin_amdgcn_readlane(x, 31));
    const float c = __int_as_float(__builtin_amdgcn_readlane(x, 47)), d = __int_as_float(__builtin_amdgcn_readlane(x, 63));
    return fminf(fminf(a, b), fminf(c, d));
}
__device__ __forceinline__ float wave_maxf(float v) {
    int x = __float_as_int(v);
    x = __float_as_int(fmaxf(__int_as_float(x), __int_as_float(dpp_shr_keep<0x111>(x))));
    x = __float_as_int(fmaxf(__int_as_float(x), __int_as_float(dpp_shr_keep<0x112>(x))));
    x = __float_as_int(fmaxf(__int_as_float(x), __int_as_float(dpp_shr_keep<0x114>(x))));
    x = __float_as_int(fmaxf(__int_as_float(x), __int_as_float(dpp_shr_keep<0x118>(x))));
    const float a = __int_as_float(__builtin_amdgcn_readlane(x, 15)), b = __int_as_float(__builtin_amdgcn_readlane(x, 31));
    const float c = __int_as_float(__builtin_amdgcn_readlane(x, 47)), d = __int_as_float(__builtin_amdgcn_readlane(x, 63));
    return fmaxf(fmaxf(a, b), fmaxf(c, d));
}

// Workgroups are dispatched to the 8 XCDs of an MI355X round robin (block b runs on XCD b % 8),
// and each XCD has its own L2. xcd_block renumbers block b of nb so that every XCD takes one
// contiguous range of logical blocks, and neighbouring work shares an L2. A bijection on [0, nb).
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
    const unsigned x = b % 8u, i = b / 8u;
    return x * (nb / 8u) + (x < nb % 8u ? x : nb % 8u) + i;
}

}  // namespace lagpm_wave
