// rocrand_xcheck.cpp -- TEST INFRASTRUCTURE ONLY.
// Host-side evaluation of rocRAND's header-only Philox4x32-10 engine
// (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h), an implementation independent
// of this project, used to pin the oracle's Philox.  engine(seed, subsequence = block,
// offset = 4 * pid) starts at counter (pid lo, pid hi, block, 0) with key
// (seed lo, seed hi), i.e. exactly block `block` of the reference stream of photon pid
// (src/philox.jl:43-57).
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdint>

extern "C" void xcheck_philox_block(uint64_t seed, uint64_t pid, uint32_t block,
                                    uint32_t out[4])
{
    rocrand_device::philox4x32_10_engine eng(seed, block, 4ull * pid);
    for (int k = 0; k < 4; ++k) out[k] = eng.next();
}
