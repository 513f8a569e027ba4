// Sub-FFT passes specialised at compile time: the table of shapes measured on MI355X for the benchmark lengths,
// and the lookup over it and the generated tables (kernels_fixed_auto*.hip).
#include <vector>

#include "fixed_impl.hpp"

namespace kbsrc::gpu {

namespace {

// Transform lengths of the benchmark workloads (48k->44.1k mono; 96k->44.1k exact and fast; 44.1k->48k).
// Several (G, T) variants may follow one another for the same (policy, n, load, store); the first is the default
// and KBSRC_TUNE="n:k,..." selects the k-th for length n (the sweep that picks the defaults).
const FixedEntry kEntries[] = {
    // exact (four-step sub-FFTs with the one-TwoSum add)
    KB_FIXED(POL_DDS, 4, 256, LD_REAL_F64, ST_CPLX, 3, 3, 5, 7),    // 315: 96k->44.1k forward columns
    KB_FIXED(POL_DDS, 8, 512, LD_REAL_F64, ST_CPLX, 3, 3, 5, 7),
    KB_FIXED(POL_DDS, 2, 128, LD_REAL_F64, ST_CPLX, 3, 3, 5, 7),
    KB_FIXED(POL_DDS, 4, 448, LD_REAL_F64, ST_CPLX, 3, 3, 5, 7),
    KB_FIXED(POL_DDS, 4, 512, LD_REAL_F64, ST_CPLX, 3, 3, 5, 7),
    KB_FIXED(POL_DDS, 7, 512, LD_REAL_F64, ST_CPLX, 3, 3, 5, 7),
    KB_FIXED(POL_DDS, 1, 64, LD_REAL_F64, ST_CPLX, 3, 3, 5, 7),
    KB_FIXED(POL_DDS, 3, 192, LD_REAL_F64, ST_CPLX, 3, 3, 5, 7),
    KB_FIXED(POL_DDS, 4, 256, LD_CPLX, ST_CPLX, 4, 4, 4, 5),        // 320: 96k->44.1k forward rows
    KB_FIXED(POL_DDS, 8, 512, LD_CPLX, ST_CPLX, 4, 4, 4, 5),
    KB_FIXED(POL_DDS, 4, 320, LD_CPLX, ST_CPLX, 4, 4, 4, 5),
    KB_FIXED(POL_DDS, 4, 512, LD_CPLX, ST_CPLX, 4, 4, 4, 5),
    KB_FIXED(POL_DDS, 2, 128, LD_CPLX, ST_CPLX, 4, 4, 4, 5),
    KB_FIXED(POL_DDS, 6, 384, LD_CPLX, ST_CPLX, 4, 4, 4, 5),
    KB_FIXED(POL_DDS, 8, 256, LD_CPLX, ST_CPLX, 3, 3, 3, 7),        // 189: 96k->44.1k inverse columns
    KB_FIXED(POL_DDS, 8, 512, LD_CPLX, ST_CPLX, 3, 3, 3, 7),
    KB_FIXED(POL_DDS, 5, 192, LD_CPLX, ST_CPLX, 3, 3, 3, 7),
    KB_FIXED(POL_DDS, 7, 320, LD_CPLX, ST_CPLX, 3, 3, 3, 7),
    KB_FIXED(POL_DDS, 7, 512, LD_CPLX, ST_CPLX, 3, 3, 3, 7),
    KB_FIXED(POL_DDS, 4, 256, LD_CPLX, ST_CPLX, 3, 3, 3, 7),
    KB_FIXED(POL_DDS, 4, 128, LD_CPLX, ST_CPLX, 3, 3, 3, 7),
    KB_FIXED(POL_DDS, 1, 64, LD_CPLX, ST_CPLX, 3, 3, 3, 7),
    KB_FIXED(POL_DDS, 5, 256, LD_CPLX, ST_OUT, 5, 7, 7),            // 245: 96k->44.1k inverse rows
    KB_FIXED(POL_DDS, 7, 512, LD_CPLX, ST_OUT, 5, 7, 7),
    KB_FIXED(POL_DDS, 9, 448, LD_CPLX, ST_OUT, 5, 7, 7),
    KB_FIXED(POL_DDS, 9, 512, LD_CPLX, ST_OUT, 5, 7, 7),
    KB_FIXED(POL_DDS, 4, 256, LD_CPLX, ST_OUT, 5, 7, 7),
    KB_FIXED(POL_DDS, 2, 128, LD_CPLX, ST_OUT, 5, 7, 7),
    KB_FIXED(POL_DDS, 10, 512, LD_CPLX, ST_OUT, 5, 7, 7),
    KB_FIXED(POL_DDS, 7, 384, LD_CPLX, ST_OUT, 5, 7, 7),
    KB_FIXED(POL_DDS, 1, 64, LD_CPLX, ST_OUT, 5, 7, 7),
    KB_FIXED(POL_DDS, 8, 256, LD_REAL_F64, ST_CPLX, 3, 3, 3, 7),    // 189: 44.1k->48k forward columns
    KB_FIXED(POL_DDS, 5, 192, LD_REAL_F64, ST_CPLX, 3, 3, 3, 7),
    KB_FIXED(POL_DDS, 5, 256, LD_CPLX, ST_CPLX, 5, 7, 7),           // 245: 44.1k->48k forward rows
    KB_FIXED(POL_DDS, 9, 448, LD_CPLX, ST_CPLX, 5, 7, 7),
    KB_FIXED(POL_DDS, 8, 448, LD_CPLX, ST_CPLX, 4, 4, 2, 7),        // 224: 44.1k->48k inverse columns
    KB_FIXED(POL_DDS, 5, 256, LD_CPLX, ST_OUT, 3, 3, 5, 5),         // 225: 44.1k->48k inverse rows
    KB_FIXED(POL_DDS, 7, 320, LD_CPLX, ST_OUT, 3, 3, 5, 5),
    KB_FIXED(POL_DDS, 5, 128, LD_REAL_F64, ST_CPLX, 5, 5, 5),       // 125: 48k->44.1k forward columns
    KB_FIXED(POL_DDS, 8, 320, LD_CPLX, ST_CPLX, 4, 4, 2, 5),        // 160: 48k->44.1k forward rows
    KB_FIXED(POL_DDS, 5, 128, LD_CPLX, ST_CPLX, 5, 5, 5),           // 125: 48k->44.1k inverse columns
    KB_FIXED(POL_DDS, 11, 320, LD_CPLX, ST_OUT, 3, 7, 7),           // 147: 48k->44.1k inverse rows
    // exact, fused: forward rows with the real-FFT split, inverse columns with x H, bin map and pre-split
    KB_FIXED(POL_DDS, 4, 320, LD_CPLX, ST_SPLIT, 4, 4, 4, 5),       // 320: 96k->44.1k
    KB_FIXED(POL_DDS, 6, 192, LD_FOLD, ST_CPLX, 3, 3, 3, 7),        // 189: 96k->44.1k
    KB_FIXED(POL_DDS, 8, 448, LD_CPLX, ST_SPLIT, 5, 7, 7),          // 245: 44.1k->48k
    KB_FIXED(POL_DDS, 8, 448, LD_FOLD, ST_CPLX, 4, 4, 2, 7),        // 224: 44.1k->48k
    KB_FIXED(POL_DDS, 8, 320, LD_CPLX, ST_SPLIT, 4, 4, 2, 5),       // 160: 48k->44.1k
    KB_FIXED(POL_DDS, 4, 128, LD_FOLD, ST_CPLX, 5, 5, 5),           // 125: 48k->44.1k
    // 336: 8-channel 384 kHz -> 44.1 kHz inverse columns and rows (ahead of the generated 4 x 256 shapes)
    KB_FIXED(POL_DDS, 6, 512, LD_CPLX, ST_CPLX, 4, 4, 3, 7),
    KB_FIXED(POL_DDS, 6, 512, LD_CPLX, ST_OUT, 4, 4, 3, 7),
    // 16 (radix 4 x 4): the plain dd baseline of the split-precision radix-16 comparison
    KB_FIXED(POL_DDS, 64, 256, LD_CPLX, ST_CPLX, 4, 4),
    KB_FIXED(POL_DDS, 32, 128, LD_CPLX, ST_CPLX, 4, 4),
    // fast (FP64)
    KB_FIXED(POL_F64, 8, 256, LD_REAL_F64, ST_CPLX, 4, 2, 5, 5),    // 200: 96k->44.1k fast forward columns
    KB_FIXED(POL_F64, 6, 128, LD_REAL_F64, ST_CPLX, 4, 2, 5, 5),
    KB_FIXED(POL_F64, 16, 320, LD_REAL_F64, ST_CPLX, 4, 2, 5, 5),
    KB_FIXED(POL_F64, 16, 640, LD_REAL_F64, ST_CPLX, 4, 2, 5, 5),
    KB_FIXED(POL_F64, 16, 640, LD_CPLX, ST_CPLX, 4, 2, 5, 5),       // 200: 96k->44.1k fast forward rows
    KB_FIXED(POL_F64, 8, 256, LD_CPLX, ST_CPLX, 4, 2, 5, 5),
    KB_FIXED(POL_F64, 6, 128, LD_CPLX, ST_CPLX, 4, 2, 5, 5),
    KB_FIXED(POL_F64, 16, 320, LD_CPLX, ST_CPLX, 4, 2, 5, 5),
    KB_FIXED(POL_F64, 20, 256, LD_CPLX, ST_CPLX, 5, 5, 5),          // 125: 96k->44.1k fast inverse columns
    KB_FIXED(POL_F64, 10, 128, LD_CPLX, ST_CPLX, 5, 5, 5),
    KB_FIXED(POL_F64, 5, 128, LD_CPLX, ST_CPLX, 5, 5, 5),
    KB_FIXED(POL_F64, 9, 256, LD_CPLX, ST_OUT, 3, 7, 7),            // 147: 96k->44.1k fast inverse rows
    KB_FIXED(POL_F64, 6, 128, LD_CPLX, ST_OUT, 3, 7, 7),
    KB_FIXED(POL_F64, 6, 128, LD_CPLX, ST_SPLIT, 4, 2, 5, 5),       // 200: 96k->44.1k fast, fused split
    KB_FIXED(POL_F64, 10, 128, LD_FOLD, ST_CPLX, 5, 5, 5),          // 125: 96k->44.1k fast, fused fold
};
#undef KB_FIXED

struct Table {
    const FixedEntry* e;
    int n;
};
std::vector<Table> tables() {   // the measured table first: its entries are the defaults where both have a shape
    std::vector<Table> t;
    t.push_back({kEntries, (int)(sizeof(kEntries) / sizeof(kEntries[0]))});
    for (auto f : {fixed_table_auto0, fixed_table_auto1, fixed_table_auto2, fixed_table_auto3, fixed_table_auto4,
                   fixed_table_extreme}) {
        int n = 0;
        const FixedEntry* e = f(&n);
        t.push_back({e, n});
    }
    return t;
}
const std::vector<Table>& all_tables() {
    static const std::vector<Table> t = tables();
    return t;
}

}  // namespace

bool has_fixed(int policy, int n, int ld, int st) { return fixed_variants(policy, n, ld, st) > 0; }

// variant index for length n from KBSRC_TUNE="n:k,n:k,..." (0 when absent)
int tune_index(int n) {
    const char* v = std::getenv("KBSRC_TUNE");
    if (!v) return 0;
    for (const char* p = v; *p;) {
        char* end = nullptr;
        const long len = std::strtol(p, &end, 10);
        if (end == p || *end != ':') break;
        const long k = std::strtol(end + 1, &end, 10);
        if (len == n) return (int)k;
        p = (*end == ',') ? end + 1 : end;
        if (*end != ',') break;
    }
    return 0;
}

int fixed_variants(int policy, int n, int ld, int st) {
    int c = 0;
    for (const Table& t : all_tables())
        for (int i = 0; i < t.n; ++i) {
            const FixedEntry& e = t.e[i];
            c += e.policy == policy && e.n == n && e.ld == ld && e.st == st;
        }
    return c;
}

bool pass_offsets_fit32(const PassArgs& a) {
    constexpr long long lim = 1LL << 27;   // elements of at most 32 bytes: byte offsets below 2^32
    const long long n = a.sd.n, m = a.nsub;
    const long long ld_max = (n - 1) * a.lde + (m - 1) * a.ldg, st_max = (n - 1) * a.ste + (m - 1) * a.stg;
    return ld_max < lim && st_max < lim && a.zh < lim && a.hopP < lim;
}

void set_prune_window(PassArgs& a, int st) {
    a.wstart = 0;
    a.wlen = -1;
    const long long n = a.sd.n;
    if (st != ST_CPLX || a.zkeep < 0 || a.ste <= 0 || n < 2) return;
    // e <= elo has a needed element at j = 0; e >= ehi at j = nsub - 1 (the last sub-FFT's offset is the largest)
    const long long elo = a.zkeep / a.ste;
    const long long top = a.zh - a.zkeep - (long long)(a.nsub - 1) * a.stg;
    const long long ehi = top <= 0 ? 0 : (top + a.ste - 1) / a.ste;
    if (elo >= n - 1 || ehi <= elo + 1) return;   // every output needed
    if (ehi >= n) {
        a.wlen = (int)(elo + 1);
    } else {
        a.wstart = (int)ehi;
        a.wlen = (int)(n - ehi + elo + 1);
    }
}

bool launch_fixed(int policy, int ld, int st, PassArgs& a, hipStream_t s, hipError_t* err, int variant) {
    if (!pass_offsets_fit32(a)) return false;   // the specialised passes index a task's elements in 32 bits
    set_prune_window(a, st);
    int skip = variant >= 0 ? variant : tune_index(a.sd.n);
    for (const Table& t : all_tables())
      for (int i = 0; i < t.n; ++i) {
        const FixedEntry& e = t.e[i];
        if (e.policy == policy && e.n == a.sd.n && e.ld == ld && e.st == st && skip-- <= 0) {
            a.G = e.G;
            if (ld == LD_FOLD || st == ST_SPLIT) a.ngrp = (a.f.M / 2 + 1 + e.G / 2 - 1) / (e.G / 2);
            else a.ngrp = (a.nsub + e.G - 1) / e.G;
            *err = e.launch(a, s);
            return true;
        }
      }
    return false;
}

}  // namespace kbsrc::gpu
