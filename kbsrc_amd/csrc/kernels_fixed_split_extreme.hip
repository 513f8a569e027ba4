// Specialised split-form passes for the extreme-ratio plans at --len 0.5 (exact): 8-channel 384 kHz -> 44.1 kHz
// (960-point forward columns from the real input, 1024-point forward rows, 336-point inverse columns and rows) and
// 10 MHz -> 48 kHz (3200-point forward rows, 100 KB of LDS, one workgroup of 768 threads per CU), at 8 and 15
// slices. Without these entries the passes run on the generic LDS split pass or through global memory.
#include "split_fixed_impl.hpp"

namespace kbsrc::gpu {

namespace {

const SplitFixedEntry kSplitExtreme[] = {
    KB_SPLIT(1, 256, LD_REAL_F64, ST_CPLX, 8, 4, 4, 4, 3, 5),
    KB_SPLIT(1, 256, LD_REAL_F64, ST_CPLX, 15, 4, 4, 4, 3, 5),
    KB_SPLIT(2, 384, LD_REAL_F64, ST_CPLX, 8, 4, 4, 4, 3, 5),
    KB_SPLIT(2, 384, LD_REAL_F64, ST_CPLX, 15, 4, 4, 4, 3, 5),
    KB_SPLIT(1, 256, LD_CPLX, ST_CPLX, 8, 4, 4, 4, 4, 4),
    KB_SPLIT(1, 256, LD_CPLX, ST_CPLX, 15, 4, 4, 4, 4, 4),
    KB_SPLIT(2, 256, LD_CPLX, ST_CPLX, 8, 4, 4, 4, 4, 4),
    KB_SPLIT(2, 256, LD_CPLX, ST_CPLX, 15, 4, 4, 4, 4, 4),
    KB_SPLIT(4, 256, LD_CPLX, ST_CPLX, 8, 4, 4, 3, 7),
    KB_SPLIT(4, 256, LD_CPLX, ST_CPLX, 15, 4, 4, 3, 7),
    KB_SPLIT(4, 256, LD_CPLX, ST_OUT, 8, 4, 4, 3, 7),
    KB_SPLIT(4, 256, LD_CPLX, ST_OUT, 15, 4, 4, 3, 7),
    KB_SPLIT(2, 128, LD_CPLX, ST_CPLX, 8, 4, 4, 3, 7),
    KB_SPLIT(2, 128, LD_CPLX, ST_CPLX, 15, 4, 4, 3, 7),
    KB_SPLIT(2, 128, LD_CPLX, ST_OUT, 8, 4, 4, 3, 7),
    KB_SPLIT(2, 128, LD_CPLX, ST_OUT, 15, 4, 4, 3, 7),
    KB_SPLIT(1, 768, LD_CPLX, ST_CPLX, 8, 4, 4, 4, 2, 5, 5),
    KB_SPLIT(1, 768, LD_CPLX, ST_CPLX, 15, 4, 4, 4, 2, 5, 5),
};

}  // namespace

const SplitFixedEntry* split_table_extreme(int* count) {
    *count = (int)(sizeof(kSplitExtreme) / sizeof(kSplitExtreme[0]));
    return kSplitExtreme;
}

}  // namespace kbsrc::gpu
