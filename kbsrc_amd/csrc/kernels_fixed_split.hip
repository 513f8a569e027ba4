// Specialised split-form passes for the sub-FFT lengths of the benchmark conversions (96 kHz -> 44.1 kHz forward
// rows, 320 points; 44.1 kHz -> 48 kHz inverse columns, 224), and the lookup over these and the extreme-ratio
// table (kernels_fixed_split_extreme.hip).
#include "split_fixed_impl.hpp"

namespace kbsrc::gpu {

namespace {

// first matching entry wins; the others stay for KBSRC_SPLIT_VARIANT sweeps. Measured on the 180 s 96 kHz ->
// 44.1 kHz conversion (ms per conversion, plain 1.767 on the same device): G = 4, T = 256: 1.752 (6 slices), 1.965
// (15); G = 2, T = 192: 1.817, 1.997; G = 4, T = 320: 1.789, 2.313.
const SplitFixedEntry kSplit[] = {
    KB_SPLIT(4, 256, LD_CPLX, ST_CPLX, 6, 4, 4, 4, 5),
    KB_SPLIT(4, 256, LD_CPLX, ST_CPLX, 7, 4, 4, 4, 5),
    KB_SPLIT(4, 256, LD_CPLX, ST_CPLX, 8, 4, 4, 4, 5),
    KB_SPLIT(4, 256, LD_CPLX, ST_CPLX, 15, 4, 4, 4, 5),
    KB_SPLIT(2, 192, LD_CPLX, ST_CPLX, 6, 4, 4, 4, 5),
    KB_SPLIT(2, 192, LD_CPLX, ST_CPLX, 7, 4, 4, 4, 5),
    KB_SPLIT(2, 192, LD_CPLX, ST_CPLX, 8, 4, 4, 4, 5),
    KB_SPLIT(2, 192, LD_CPLX, ST_CPLX, 15, 4, 4, 4, 5),
    KB_SPLIT(4, 320, LD_CPLX, ST_CPLX, 6, 4, 4, 4, 5),
    KB_SPLIT(4, 320, LD_CPLX, ST_CPLX, 15, 4, 4, 4, 5),
    KB_SPLIT(4, 256, LD_CPLX, ST_CPLX, 6, 4, 4, 2, 7),
    KB_SPLIT(4, 256, LD_CPLX, ST_CPLX, 7, 4, 4, 2, 7),
    KB_SPLIT(4, 256, LD_CPLX, ST_CPLX, 8, 4, 4, 2, 7),
    KB_SPLIT(4, 256, LD_CPLX, ST_CPLX, 15, 4, 4, 2, 7),
};

}  // namespace

const SplitFixedEntry* find_split_fixed(int n, int ld, int st, int slices) {
    const char* v = std::getenv("KBSRC_SPLIT_VARIANT");   // sweep: take the k-th matching entry
    int skip = v ? std::atoi(v) : 0;
    int next = 0;
    const SplitFixedEntry* more = split_table_extreme(&next);
    for (const SplitFixedEntry& e : kSplit)
        if (e.n == n && e.ld == ld && e.st == st && e.slices == slices && skip-- <= 0) return &e;
    for (int i = 0; i < next; ++i) {
        const SplitFixedEntry& e = more[i];
        if (e.n == n && e.ld == ld && e.st == st && e.slices == slices && skip-- <= 0) return &e;
    }
    return nullptr;
}

}  // namespace kbsrc::gpu
