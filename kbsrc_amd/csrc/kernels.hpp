// gfx950 kernels of the overlap-save engine (declarations of their argument blocks and launchers).
#pragma once
#include <hip/hip_runtime.h>

#include "dd.hpp"

namespace kbsrc::gpu {

constexpr int kMaxStages = 24;

// One Stockham stage of a sub-FFT: radix r, m = (remaining length) / r, s = product of earlier radices, and the
// offset (in table elements) of its twiddles w[p*(r-1) + u-1] = cis(-2 pi s p u / n).
struct StageDesc {
    int r, m, s, tw;
};
struct SubDesc {
    int n = 0, nst = 0;
    StageDesc st[kMaxStages];
};

enum LoadKind { LD_REAL_F64 = 0, LD_REAL_DD = 1, LD_CPLX = 2, LD_FOLD = 3 };
enum StoreKind { ST_CPLX = 0, ST_OUT = 1, ST_SPLIT = 2 };

// Real-FFT split, spectrum product and bin map fused into the passes next to them (specialised passes only):
//   ST_SPLIT  forward row pass: stores X[i] = forward split of z (src/dsp/fft.hpp:365-407) for i <= xmax;
//   LD_FOLD   inverse column pass: loads X, forms Y = X H with the bin map (src/resampler.cpp:64-78) and the
//             inverse pre-split (src/dsp/fft.hpp:412-458) before its sub-FFTs.
// Both need each sub-FFT j in the same workgroup as its mirror M - j: the workgroup takes slot pairs
// {s, M - s} (slot 0 and, for even M, slot M/2 alone).
struct FuseArgs {
    const double* H = nullptr;      // real kernel spectrum, (hi, lo) pairs or FP64
    const double* wf = nullptr;     // forward split twiddles w[0..h]
    const double* wi = nullptr;     // inverse split twiddles w[0..hinv]
    int h = 0, hinv = 0, Nb = 0, Mb = 0, K = 0, jmax = 0, up = 0, xmax = 0;
    int M = 0;                      // mirror modulus of the sub-FFT index: N1 (rows) or N2 (columns)
};
enum Policy { POL_F64 = 0, POL_DDA = 1, POL_DDS = 2 };

// A batch of equal sub-FFTs, G of them per workgroup, staged through LDS. Element e of sub-FFT j (within a
// task) is read at complex index e*lde + j*ldg and written at e*ste + j*stg.
// Where (channel, block) task gtask reads its input block and writes its output samples. One job: channel
// gtask / nblk, block gtask % nblk, input at channel * in_cstride + block * in_bstride, output channel at
// channel * nout. A batch of jobs gives a table instead: tdesc[4 gtask + 0..3] = the block's input offset, the
// channel's output offset (both in doubles), the block index and the channel's output length.
struct TaskIO {
    long long in_off, out_off, blk, nout;
};
__host__ __device__ inline TaskIO task_io(const long long* tdesc, long long gtask, long long nblk, long long in_cstride,
                                          long long in_bstride, long long nout) {
    if (tdesc) return {tdesc[4 * gtask], tdesc[4 * gtask + 1], tdesc[4 * gtask + 2], tdesc[4 * gtask + 3]};
    const long long ch = gtask / nblk, blk = gtask - ch * nblk;
    return {ch * in_cstride + blk * in_bstride, ch * nout, blk, nout};
}

struct PassArgs {
    SubDesc sd;
    const double* tw = nullptr;     // stage twiddles of this sub-FFT length
    const double* oddc = nullptr;   // radix 3/5/7 constants: element (ri*7 + k) = cis(2 pi k / R)
    int G = 1, nsub = 1, ngrp = 1, ntask = 1;
    long long task0 = 0;            // global (channel, block) task number of local task 0
    // load side
    const double* src = nullptr;
    long long src_tstride = 0;      // doubles between tasks (LD_CPLX, LD_REAL_DD)
    long long lde = 1, ldg = 1;
    long long nblk = 1, in_cstride = 0, in_bstride = 0;   // LD_REAL_F64: task t reads in + c*cstride + b*bstride
    const long long* tdesc = nullptr;   // a batch of jobs: per-task addresses instead (see task_io)
    // store side
    double* dst = nullptr;
    long long dst_tstride = 0;
    long long ste = 1, stg = 1;
    const double* tw4 = nullptr;    // four-step twiddles, indexed like the store address
    int twmode = 0;                 // 0: none, 1: every e != 0, 2: e != 0 and sub-FFT index != 0
    int renorm = 0;                 // QuickTwoSum each stored dd component
    // ST_OUT: unpack r[2i] = re, r[2i+1] = -im, keep [margP, margP + hopP), write out[c*nout + b*hopP + x - o]
    double* out = nullptr;
    long long nout = 0, o = 0, margP = 0, hopP = 0;
    int out_dd = 0;                 // ST_OUT, dd policies: each sample as its (hi, lo) pair, at 2 (c*nout + d)
    FuseArgs f;
    double* gscr0 = nullptr;        // global scratch for sub-FFTs too long for LDS: ntask * nsub * n elements each
    double* gscr1 = nullptr;
    // pm > 0: the pass runs the first stages of longer transforms of length N = pm * sd.n, one per (task, sub-FFT)
    // column c = the sub-FFT index; the butterfly with index p here is butterfly c + pm * p of the long transform
    // and takes that one's stage twiddle from sd's tables (indexed for length N).
    int pm = 0;
    // split > 0 (exact-mode plans only): the split-precision form. A sub-FFT whose first two stages are radix 4
    // runs them as one radix-16 int8-MFMA stage with `split` base-128 slices per operand (adig: the DFT matrix's
    // slices in the MFMA A layout; tw16: w_n^{p k} at p * 15 + k - 1, p < n / 16, k = 1..15), through gscr0/1.
    int split = 0;
    const double* tw16 = nullptr;
    const long* adig = nullptr;
    // zkeep >= 0 (ST_CPLX, the last pass of a forward transform of length zh): outputs zkeep < i < zh - zkeep are
    // never read (the bin map uses the spectrum only up to bin zkeep, and the real-FFT split pairs bin i with
    // zh - i), so they are not stored
    long long zkeep = -1, zh = 0;
    // The sub-FFT outputs e that some reader needs under zkeep (set_prune_window): the cyclic window
    // [wstart, wstart + wlen) mod sd.n, wlen < 0 for all. The specialised passes run only the butterflies whose
    // outputs reach it.
    int wstart = 0, wlen = -1;
    // > 0: workgroups take the tasks in groups of tgroup, each group's sub-FFT groups in turn with its tasks fastest
    // (task_group_order), so that the tasks of one sub-FFT group, which read the same twiddles (and, for the
    // forward columns, overlapping input), run back to back on one XCD; 0: tasks slowest
    int tgroup = 0;
};
// true when a pass with zkeep set may skip storing output element i
__host__ __device__ inline bool z_unused(const PassArgs& a, long long i) { return a.zkeep >= 0 && i > a.zkeep && i < a.zh - a.zkeep; }

// Spectrum x H, bin map and inverse pre-split for one batch of tasks (src/fft.hpp:361-409,411-458;
// src/resampler.cpp:53-78): z (forward complex FFT output, h per task) -> zi (inverse complex FFT input, hi).
struct MidArgs {
    const double* z = nullptr;
    long long z_tstride = 0;
    double* zi = nullptr;
    long long zi_tstride = 0;
    const double* H = nullptr;      // real kernel spectrum
    const double* wf = nullptr;     // forward split twiddles w[0..h]
    const double* wi = nullptr;     // inverse split twiddles w[0..hi]
    int h = 0, hinv = 0, Nb = 0, Mb = 0, K = 0, jmax = 0, up = 0, ntask = 0;
    int Jg = 0;                     // > 0: general fold of H[0..Jg] on the finer kernel grid (near-unity ratios)
    int kchain = 0;                 // > 0: bin blocks kchain apart run one after another (mid_block)
    int tgroup = 0;                 // > 0: tasks in groups of tgroup, each group's bin blocks in turn (mid_order)
    // H, wf and wi residue-major for the fused row pass (entry i at (i mod t) Lr + i / t, t = gcd(N1, hinv)), or null
    const double* Hr = nullptr;
    const double* wfr = nullptr;
    const double* wir = nullptr;
    int Lr = 0;
};

// Bluestein (chirp-z) steps of a complex FFT whose length n has a prime factor above 7 (src/dsp/fft.hpp:228-278),
// around an inner 7-smooth FFT of length m >= 2n - 1:
//   pre   u[j] = x[j] conj(b[j]) for j < n, 0 up to m        (x read like a pass's load side)
//   mid   U[k] = conj(U[k] Vf[k])                           (in place)
//   post  y[k] = conj(C[k]) conj(b[k]) for k < n             (stored like a pass's store side)
struct BlueArgs {
    int n = 0, m = 0, ntask = 0;
    long long task0 = 0, nblk = 1, in_cstride = 0, in_bstride = 0;
    const long long* tdesc = nullptr;
    const double* src = nullptr;
    long long src_tstride = 0;
    double* dst = nullptr;
    long long dst_tstride = 0;
    const double* bc = nullptr;     // chirp b[t] = cis(t^2 mod 2n, 2n), n elements
    const double* vf = nullptr;     // inner FFT of the chirp sequence / m, m elements
    double* out = nullptr;          // ST_OUT, as PassArgs
    long long nout = 0, o = 0, margP = 0, hopP = 0;
    int out_dd = 0;
};
enum BluePhase { BLUE_PRE = 0, BLUE_MID = 1, BLUE_POST = 2 };

// Zero-padded planar staging of interleaved input with the edge extensions (src/resampler.cpp:168-186).
struct StageArgs {
    const double* in = nullptr;     // interleaved frames x ch
    const double* head = nullptr;   // ch x ext: sample -1-i
    const double* tail = nullptr;   // ch x ext: sample nin+i
    double* xx = nullptr;           // ch x stride, the first tot of each channel written
    long long nin = 0, tot = 0, stride = 0, off = 0, ext = 0;
    int ch = 0;
};

// Forward real-FFT split of the kernel spectrum and division by the scale (src/design.cpp:46-48).
struct SpecArgs {
    const double* z = nullptr;
    const double* w = nullptr;
    double* H = nullptr;            // (hi, lo) pairs, h+1 of them
    int h = 0;
    double scale = 1;
};

// Kaiser-sinc taps h[0..L] on the GPU (src/design.cpp:17-37): ratio = a/b and i0 = I0(beta) as dd, h as
// (hi, lo) pairs.
struct TapArgs {
    long long L = 0, a = 0, b = 1;
    double beta = 0, ratio_hi = 0, ratio_lo = 0, i0_hi = 0, i0_lo = 0;
    double* h = nullptr;
};

// Output packing (src/io/wav.cpp:52-86,235-250): frames [frame0, frame0 + frames) of the planar output y (channel
// stride nout) -> interleaved sample bytes at out (bps bytes each, set by the launcher).
enum PackFormat { PACK_F64 = 0, PACK_F32 = 1, PACK_S32 = 2, PACK_S24 = 3, PACK_S16 = 4 };
struct PackArgs {
    const double* y = nullptr;
    long long nout = 0, frame0 = 0, frames = 0;
    int ch = 1, format = PACK_F64, bps = 8;
    unsigned char* out = nullptr;
    unsigned long long* clipped = nullptr;   // incremented once per clipped sample
};
int pack_bytes(int format);   // bytes per sample, 0 for an unknown format

// Twiddle tables on the GPU from a CisTable's two short tables (a: cis(jB, N), b: cis(l, N), 4 doubles each):
// mode TAB_FOUR: element k1*N2 + n2 = cis(-k1 n2, N) for count = N entries (the four-step table); mode TAB_LINEAR:
// element i = cis(-i, N), i < count (the real-FFT split table). Elements in the plan's format (as_dd: {re.hi,
// im.hi, re.lo, im.lo}, else {re, im}).
enum TableMode { TAB_FOUR = 0, TAB_LINEAR = 1 };
struct TableArgs {
    const double* a = nullptr;
    const double* b = nullptr;
    long long N = 1, B = 1, count = 0, N2 = 1;
    int mode = TAB_LINEAR, as_dd = 1;
    double* out = nullptr;
};

size_t pass_lds_bytes(int policy, int n, int G);
hipError_t launch_pass(int policy, int ld, int st, const PassArgs& a, int threads, hipStream_t s);
// Compile-time specialised pass for the benchmark transform lengths; false when (policy, n, ld, st) has none.
// Sets a.G and a.ngrp to the specialisation's group size.
// A specialised pass: policy, length, load and store kinds, sub-FFTs per workgroup G, threads T, and its launcher.
struct FixedEntry {
    int policy, n, ld, st, G, T;
    hipError_t (*launch)(const PassArgs&, hipStream_t);
};
// generated tables of compile-time shapes (kernels_fixed_auto*.hip)
const FixedEntry* fixed_table_auto0(int* count);
const FixedEntry* fixed_table_auto1(int* count);
const FixedEntry* fixed_table_auto2(int* count);
const FixedEntry* fixed_table_auto3(int* count);
const FixedEntry* fixed_table_auto4(int* count);
const FixedEntry* fixed_table_extreme(int* count);
// specialised split-form passes (kernels_fixed_split.hip): exact-mode (POL_DDS) sub-FFT passes with the first two
// radix-4 stages as the int8-MFMA radix-16 stage; nullptr when none is built for (n, ld, st, slices)
struct SplitFixedEntry {
    int n, ld, st, slices, G, T;
    hipError_t (*launch)(const PassArgs&, hipStream_t);
};
const SplitFixedEntry* find_split_fixed(int n, int ld, int st, int slices);
const SplitFixedEntry* split_table_extreme(int* count);   // kernels_fixed_split_extreme.hip

// true when every element index of one task in pass a (loads, stores, four-step twiddles) is below 2^27: the
// specialised passes (plain and split) address a task's elements with 32-bit offsets
bool pass_offsets_fit32(const PassArgs& a);
// Sets a.wstart, a.wlen from a.zkeep for a pass with store kind st: element e of sub-FFT j goes to
// i = e ste + j stg, needed when i <= zkeep or i >= zh - zkeep for some j < nsub.
void set_prune_window(PassArgs& a, int st);
// variant: index among the registered shapes of (policy, n, ld, st); -1: the default (or KBSRC_TUNE's)
bool launch_fixed(int policy, int ld, int st, PassArgs& a, hipStream_t s, hipError_t* err, int variant = -1);
int fixed_variants(int policy, int n, int ld, int st);
bool has_fixed(int policy, int n, int ld, int st);
// A long sub-FFT pass as two specialised passes (kernels_fixed_long.hip): the leading stages on its columns, the
// rest on its classes, through scr (scr_ts doubles per task, at least n * nsub elements); same bits as one pass.
bool long_two_part(int policy, int n, int ld, int st);
hipError_t launch_long_two_part(int policy, int ld, int st, const PassArgs& a, double* scr, long long scr_ts,
                                hipStream_t s);
// Sub-FFT pass through global memory, one launch per stage (lengths whose tile does not fit LDS).
hipError_t launch_global(int policy, int ld, int st, const PassArgs& a, hipStream_t s);
hipError_t launch_mid(int policy, const MidArgs& a, hipStream_t s);
// The forward row pass and the split/multiply/fold step as one launch (kernels_rowmid.hip), for a row pass over n1
// rows of n2 points (no general fold): a registered shape exists for (policy, n2, n1, hinv). a: the
// row pass (its load side and stage tables; the prune window set), m: the split/fold step (zi, its destination).
bool rowmid_available(int policy, int n2, int n1, int hinv);
hipError_t launch_rowmid(int policy, const PassArgs& a, const MidArgs& m, hipStream_t s);
// Its split-precision form (kernels_rowmid_split.hip): a row pass of the split form (a.split slices, a.adig,
// a.tw16 set), the rows' first two radix-4 stages as the int8-MFMA radix-16 stage.
bool rowmid_split_available(int n2, int n1, int hinv, int slices);
// The plan shapes above whose fused kernels read the residue-major tables (MidArgs::Hr, wfr, wir; t = gcd(n1, hinv)).
bool rowmid_residue_tables(int policy, int n2, int n1, int hinv);
// dst[((i mod t) L + i / t) e + c] = src[i e + c] for i < n, c < e (doubles per entry)
hipError_t launch_residue_major(const double* src, long long n, int e, int t, int L, double* dst, hipStream_t s);
hipError_t launch_rowmid_split(const PassArgs& a, const MidArgs& m, hipStream_t s);
// One launch over several plan shapes' passes (device arrays: descriptors d[0..n), workgroup starts[0..n) with shape s
// taking [starts[s], starts[s + 1]) of total): the generic sub-FFT pass (each d[s] with its G, ngrp set for 256
// threads; lds = the largest tile) and the split/multiply/fold kernel (mid_workgroups each).
hipError_t launch_pass_multi(int policy, int ld, int st, const PassArgs* d, const int* starts, int n, int total,
                             size_t lds, hipStream_t s);
hipError_t launch_mid_multi(int policy, const MidArgs* d, const int* starts, int n, int total, hipStream_t s);
int mid_workgroups(const MidArgs& a);   // workgroups of launch_mid for a, -1 if too many
hipError_t launch_stage(const StageArgs& a, hipStream_t s);
// The plain dd radix-16 stage of nvec 16-point vectors with its stage exchange inside the wave (kernels_inwave.hip):
// tw and oddc the 16-point sub-FFT's stage twiddles and the radix constants (dd format). Same bits as the LDS pass.
hipError_t launch_dft16_inwave(const double* x, long long nvec, double* y, const double* tw, const double* oddc,
                               hipStream_t s);
// policy: POL_DDA or POL_F64; kind: the LD kind (pre) or ST kind (post)
hipError_t launch_blue(int policy, int phase, int kind, const BlueArgs& a, hipStream_t s);
hipError_t launch_spec_split(const SpecArgs& a, hipStream_t s);
hipError_t launch_taps(const TapArgs& a, hipStream_t s);
// circular placement of taps h[0..Lk] (pairs) in length n (src/design.cpp:42-43), pairs out
hipError_t launch_circ(const double* h, long long Lk, long long n, double* r, hipStream_t s);
// hi parts of n (hi, lo) pairs
hipError_t launch_hi(const double* pairs, long long n, double* out, hipStream_t s);
hipError_t launch_pack(PackArgs a, hipStream_t s);
// Split-precision radix-16 DFT (int8 MFMA slices; not on the bit-exact path): nvec vectors of 16 complex dd (plan
// element format) x -> y; adig: the S slices of M / 2 in the MFMA A layout, (slice, M-tile, lane) -> 8 bytes.
bool split_slices_ok(int slices);   // the slice counts built: 4-8, 10, 12, 14, 15, 16
hipError_t launch_split_dft16(const double* x, long long nvec, int slices, const long* adig, double* y,
                              hipStream_t s);
// The same as the first stage of nsub sub-FFTs of length n (16 | n) stored one after another in x: y[t n + 16 p + k]
// = w_n^{p k} DFT16(x[t n + p + (n / 16) m])[k] (tw16 as in PassArgs), dd elements.
hipError_t launch_split16_stage(const double* x, double* y, int n, long long nsub, int slices, const double* tw16,
                                const long* adig, hipStream_t s);
// A sub-FFT pass in the split form (PassArgs::split): gather to gscr0, the radix-16 split stage, the remaining
// stages as one LDS pass per class q + 16 j, scatter (four-step twiddle, renormalisation, output as the pass does).
hipError_t launch_split_pass(int policy, int ld, int st, const PassArgs& a, hipStream_t s);
// The same inside one LDS pass of the generic kernel (G sub-FFTs per workgroup as a's G, ngrp): exact-mode
// four-step sub-FFTs (POL_DDS) that fit a tile, slice counts 6, 7, 8 and 15; split_lds_ok says whether a pass
// qualifies. Same bits as the global-memory route.
bool split_lds_ok(int policy, int ld, int st, const PassArgs& a);
hipError_t launch_split_lds(int policy, int ld, int st, const PassArgs& a, int threads, hipStream_t s);
hipError_t launch_cis_table(const TableArgs& a, hipStream_t s);
// CisTable(N)'s short tables on the GPU (block B = ceil(sqrt(N)), J = N / B + 1 entries in a, B in b; 4 doubles each)
hipError_t launch_cis_short(long long N, long long B, long long J, double* a, double* b, hipStream_t s);
// Bluestein chirp b[t] = cis(t^2 mod 2n, 2n), t < n, into bc (plan format) and the chirp sequence v (dd format,
// m elements, zeroed by the caller): v[t] = v[m - t] = b[t] (src/dsp/fft.hpp:231-236)
hipError_t launch_chirp(long long n, long long m, int as_dd, double* bc, double* v, hipStream_t s);
// out[q] = V[q] / m per dd component (V dd format, m elements), out in the plan's format (src/dsp/fft.hpp:237-238)
hipError_t launch_div_m(const double* V, long long m, int as_dd, double* out, hipStream_t s);

}  // namespace kbsrc::gpu
