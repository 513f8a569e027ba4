// Specialised sub-FFT passes for the extreme-ratio plans: 8-channel 384 kHz -> 44.1 kHz and 10 MHz -> 48 kHz at
// --len 0.5, both modes (exact: 960 x 1024 and 3150 x 3200 forward four-steps; fast: 640 x 640 and 3125 x 3240).
// The 3150- and 3200-point exact sub-FFTs take 100 KB of LDS each, one per workgroup, 1024 threads (no thread
// holds more than 8 dd values across a stage); without these entries they run through global memory.
#include "fixed_impl.hpp"

namespace kbsrc::gpu {

namespace {

const FixedEntry kExtreme[] = {
    // 8-channel 384 kHz -> 44.1 kHz exact: 2 x 512 measured ahead of the generated 1 x 256 and 2 x 256 (second)
    KB_FIXED(POL_DDS, 2, 512, LD_REAL_F64, ST_CPLX, 4, 4, 4, 3, 5),
    KB_AUTO(POL_DDS, 960, LD_REAL_F64, ST_CPLX),
    KB_FIXED(POL_DDS, 2, 512, LD_CPLX, ST_CPLX, 4, 4, 4, 4, 4),
    KB_AUTO(POL_DDS, 1024, LD_CPLX, ST_CPLX),
    // 3150 and 3200 points, 16 waves per workgroup (10 waves, T = 640, the second entries: 28.6 against 27.5 ms
    // for the 30 s 10 MHz conversion)
    KB_FIXED(POL_DDS, 1, 1024, LD_REAL_F64, ST_CPLX, 2, 3, 3, 5, 5, 7),
    KB_FIXED(POL_DDS, 1, 1024, LD_CPLX, ST_CPLX, 4, 4, 4, 2, 5, 5),
    KB_FIXED(POL_DDS, 1, 640, LD_REAL_F64, ST_CPLX, 2, 3, 3, 5, 5, 7),
    KB_FIXED(POL_DDS, 1, 640, LD_CPLX, ST_CPLX, 4, 4, 4, 2, 5, 5),
    KB_AUTO(POL_F64, 960, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_F64, 1024, LD_CPLX, ST_CPLX),
    KB_AUTO(POL_F64, 640, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_F64, 640, LD_CPLX, ST_CPLX),
    KB_AUTO(POL_F64, 224, LD_CPLX, ST_OUT),
    KB_AUTO(POL_F64, 225, LD_CPLX, ST_OUT),
    KB_AUTO(POL_F64, 336, LD_CPLX, ST_CPLX),
    KB_AUTO(POL_F64, 336, LD_CPLX, ST_OUT),
    KB_AUTO(POL_F64, 3125, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_F64, 3240, LD_CPLX, ST_CPLX),
};

}  // namespace

const FixedEntry* fixed_table_extreme(int* count) {
    *count = (int)(sizeof(kExtreme) / sizeof(kExtreme[0]));
    return kExtreme;
}

}  // namespace kbsrc::gpu
