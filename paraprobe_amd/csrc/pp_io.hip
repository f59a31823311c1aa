// pp_io.hip -- gfx950 kernel that decodes IVAS POS records and ranges them.
//
// Replaces solverHdl::load_pos_sequentially's per-record conversion
// (src/PARAPROBE_SolverHdl.cpp:6797-6930; pos(pos_be_2_stl_le),
// src/PARAPROBE_EPOSEndianness.h:45-84: four big-endian IEEE-754 floats x, y,
// z, mass-to-charge per 16-byte record, byte-reversed into host order) and
// solverHdl::identify_ions / PeriodicTable::mq2ionname
// (S.cpp:7340-7438, src/PARAPROBE_PeriodicTable.cpp:65-91): the type of an
// ion is the type of the first range, in the caller's order (type id
// ascending, Unknown's dummy range [0, 0.001] first), whose closed interval
// holds mq (`!(mq < lo) && !(mq > hi)`, so a NaN mq matches the first range);
// no match gives type 0 (UNKNOWNTYPE).  One thread per record; the ranges sit
// in LDS.  HBM-bound: 16 bytes read (POS; 44 bytes of EPOS records, of which
// 16 are used) and 17 written per ion.
#include "pp_launch.h"

namespace pp {

namespace {

inline unsigned blocks_for(uint64_t n, unsigned block)
{
    uint64_t g = (n + block - 1) / block;
    if (g > 0x7FFFFFFFull)
        g = 0x7FFFFFFFull;
    return (unsigned)(g ? g : 1);
}

__device__ inline float be_float(uint32_t w) { return __uint_as_float(__builtin_bswap32(w)); }

} // namespace

// W: 32-bit words per record, 4 for POS (x, y, z, mq) and 11 for EPOS (x, y,
// z, mq, tof, vdc, vpu, detx, dety and two ints, src/PARAPROBE_EPOSEndianness.h:
// 86-160); the first four words are the ones decoded here.
template <int W>
__global__ __launch_bounds__(256) void k_posdecode(const uint32_t *__restrict__ rec, uint64_t n,
                                                   const float *__restrict__ lo,
                                                   const float *__restrict__ hi,
                                                   const uint8_t *__restrict__ rtype, int nranges,
                                                   float *__restrict__ xyz,
                                                   float *__restrict__ mq,
                                                   uint8_t *__restrict__ type)
{
    extern __shared__ float s_rng[];  // lo[nranges], hi[nranges], type bytes
    float *s_lo = s_rng, *s_hi = s_rng + nranges;
    uint8_t *s_t = reinterpret_cast<uint8_t *>(s_rng + 2 * nranges);
    for (int r = threadIdx.x; r < nranges; r += blockDim.x) {
        s_lo[r] = lo[r];
        s_hi[r] = hi[r];
        s_t[r] = rtype[r];
    }
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint4 w;
    if (W == 4) {
        w = reinterpret_cast<const uint4 *>(rec)[i];
    } else {
        const uint32_t *r = rec + (uint64_t)W * i;
        w = make_uint4(r[0], r[1], r[2], r[3]);
    }
    const float x = be_float(w.x), y = be_float(w.y), z = be_float(w.z), m = be_float(w.w);
    uint8_t t = 0;
    for (int r = 0; r < nranges; ++r)
        if (!(m < s_lo[r]) && !(m > s_hi[r])) {
            t = s_t[r];
            break;
        }
    xyz[3 * i] = x;
    xyz[3 * i + 1] = y;
    xyz[3 * i + 2] = z;
    if (mq)
        mq[i] = m;
    type[i] = t;
}

hipError_t launch_posdecode(const uint32_t *rec, uint64_t n, int words, const float *lo,
                            const float *hi, const uint8_t *rtype, int nranges, float *xyz,
                            float *mq, uint8_t *type, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    const size_t lds = (size_t)nranges * 9 + 16;
    if (words == 4)
        hipLaunchKernelGGL(k_posdecode<4>, dim3(blocks_for(n, 256)), dim3(256), lds, s, rec, n,
                           lo, hi, rtype, nranges, xyz, mq, type);
    else if (words == 11)
        hipLaunchKernelGGL(k_posdecode<11>, dim3(blocks_for(n, 256)), dim3(256), lds, s, rec, n,
                           lo, hi, rtype, nranges, xyz, mq, type);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

} // namespace pp
