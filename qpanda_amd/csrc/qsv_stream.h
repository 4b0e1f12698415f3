// qsv_stream.h -- helpers of the streaming sweeps (device code only).
//
// Nontemporal accesses for sweeps that read and write each amplitude once,
// and the XCD block mapping: workgroup b of a launch runs on XCD b % 8
// (round-robin dispatch), so xcd_block() gives each XCD runs of 64
// consecutive blocks of the index range (grids that are not a multiple of 512
// keep b).  Measured at 30 qubits, bit-identical: window sweeps 2-5% faster,
// the uncontrolled 2x2 gate sweep 6.40 -> 6.11 ms, gate-tiled sweeps 4.5%,
// single-block fused sweeps 6%, the FP64 norm of a c128 state 2.88 -> 2.46
// ms, <a|b> 11-13%, marginals 7%, collapse 10% (the reductions keep their
// partition and summation order); sweeps that touch only part of each cache
// line (controlled gates) ran slower with nontemporal accesses and keep
// cached ones.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qsv {

template <typename T> struct NtVec;
template <> struct NtVec<double2> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct NtVec<float2> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct NtVec<float4> { typedef float type __attribute__((ext_vector_type(4))); };

// (component copies: a bit cast of double2 made the 2x2 gate sweep 4% slower)
__device__ __forceinline__ double2 ld_nt(const double2 *p)
{
    const NtVec<double2>::type v = __builtin_nontemporal_load(reinterpret_cast<const NtVec<double2>::type *>(p));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ float2 ld_nt(const float2 *p)
{
    const NtVec<float2>::type v = __builtin_nontemporal_load(reinterpret_cast<const NtVec<float2>::type *>(p));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float4 ld_nt(const float4 *p)
{
    const NtVec<float4>::type v = __builtin_nontemporal_load(reinterpret_cast<const NtVec<float4>::type *>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st_nt(double2 *p, double2 x)
{
    const NtVec<double2>::type v = {x.x, x.y};
    __builtin_nontemporal_store(v, reinterpret_cast<NtVec<double2>::type *>(p));
}
__device__ __forceinline__ void st_nt(float2 *p, float2 x)
{
    const NtVec<float2>::type v = {x.x, x.y};
    __builtin_nontemporal_store(v, reinterpret_cast<NtVec<float2>::type *>(p));
}
__device__ __forceinline__ void st_nt(float4 *p, float4 x)
{
    const NtVec<float4>::type v = {x.x, x.y, x.z, x.w};
    __builtin_nontemporal_store(v, reinterpret_cast<NtVec<float4>::type *>(p));
}

__device__ __forceinline__ uint64_t xcd_block(uint64_t b, uint64_t grid)
{
    if (grid & 511) return b;
    const uint64_t q = b >> 3;
    return (q >> 6) * 512 + (b & 7) * 64 + (q & 63);
}

}  // namespace qsv
