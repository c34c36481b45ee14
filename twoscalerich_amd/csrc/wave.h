// wave.h - wavefront reductions shared by the HIP kernels (device code only).
#pragma once
#include <hip/hip_runtime.h>

namespace tsr {

// Device arrays reached through pointers held in memory (basis column tables, the problem
// description's structs) are generic pointers to the compiler.  Reading them through
// global address-space pointers lets it issue global rather than flat loads: a flat
// load may alias LDS, so it is ordered against the kernels' LDS accesses and waited for
// one at a time, while global loads stay in flight together.
template <class T>
using gptr = const __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> as_global(const T* p)
{
    return (gptr<T>)p;
}
template <class T>
using gptr_w = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr_w<T> as_global_w(T* p)
{
    return (gptr_w<T>)p;
}

// x of lanes the DPP control (row/bank masks, out-of-row sources) leaves out read as 0
template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ double dpp_or_zero(double x)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROW_MASK, BANK_MASK, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROW_MASK, BANK_MASK, false);
    return __hiloint2double(hi, lo);
}

// Wavefront sum in a fixed order with DPP lane moves (VALU, no LDS traffic; every lane of
// the wavefront must be active): pairs and quads by quad permutes, the rest of each
// 16-lane row by row shifts, then rows 0+1 and 2+3 by row broadcasts; lane 63 holds the
// total, read back to every lane.
__device__ __forceinline__ double wave_sum_dpp(double x)
{
    x += dpp_or_zero<0xB1, 0xf, 0xf>(x);    // quad_perm [1,0,3,2]
    x += dpp_or_zero<0x4E, 0xf, 0xf>(x);    // quad_perm [2,3,0,1]
    x += dpp_or_zero<0x114, 0xf, 0xe>(x);   // row_shr:4 into lanes 4..15 of each row
    x += dpp_or_zero<0x118, 0xf, 0xc>(x);   // row_shr:8 into lanes 8..15: lane 15 = row sum
    x += dpp_or_zero<0x142, 0xa, 0xf>(x);   // row_bcast:15 into rows 1, 3
    x += dpp_or_zero<0x143, 0xc, 0xf>(x);   // row_bcast:31 into rows 2, 3: lane 63 = total
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), 63);
    return __hiloint2double(hi, lo);
}

}  // namespace tsr
