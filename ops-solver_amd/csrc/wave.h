// Wave64 reductions on gfx950 shared by the real LU kernels: a 32-bit unsigned maximum over
// the 64 lanes (DPP row operations inside each row of 16 lanes, then the four row results
// read with v_readlane), and the 64-bit maximum of (magnitude, position) pivot keys reduced
// word by word.
#pragma once
#include <hip/hip_runtime.h>

namespace ops
{
namespace
{
template <int CTRL>
__device__ inline unsigned DppU32(unsigned v)
{
    return static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, 0xF, 0xF, false));
}
__device__ inline unsigned WaveMaxU32(unsigned v)
{
    auto mx = [](unsigned a, unsigned b) { return a > b ? a : b; };
    v = mx(v, DppU32<0xB1>(v));   // quad_perm [1,0,3,2]
    v = mx(v, DppU32<0x4E>(v));   // quad_perm [2,3,0,1]
    v = mx(v, DppU32<0x124>(v));  // row_ror:4
    v = mx(v, DppU32<0x128>(v));  // row_ror:8
    const unsigned r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const unsigned r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return mx(mx(r0, r1), mx(r2, r3));
}
// The maximum of a 64-bit key over the wave, reduced word by word: the high words first
// (one 32-bit DPP max per step), then, only when several lanes hold that high word, their
// low words.
// Keys here hold a magnitude in the high word and a position in the low one, so the
// second pass runs only on exact magnitude ties.
__device__ inline unsigned long long WaveMaxU64Split(unsigned long long v)
{
    const unsigned hi = static_cast<unsigned>(v >> 32), lo = static_cast<unsigned>(v);
    const unsigned mh = WaveMaxU32(hi);
    const unsigned long long tie = __ballot(hi == mh);
    unsigned ml;
    if (__popcll(tie) == 1)  // uniform
        ml = __builtin_amdgcn_readlane(lo, __builtin_ctzll(tie));
    else
        ml = WaveMaxU32(hi == mh ? lo : 0u);
    return (static_cast<unsigned long long>(mh) << 32) | ml;
}
}  // namespace
}  // namespace ops
