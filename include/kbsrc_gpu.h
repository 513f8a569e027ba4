/* KBSRC on MI355X: the C-ABI boundary of the GPU resampler (libkbsrc_gpu.so).
 *
 * Plain C types only; every call returns a status and never throws. The C++ API in include/kbsrc/kbsrc.hpp is
 * a thin wrapper over these entry points. Each declaration names the reference interface it replaces.
 */
#ifndef KBSRC_GPU_H
#define KBSRC_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. EINVAL carries the reference's std::invalid_argument messages (src/resampler.cpp:94-155). */
#define KBSRC_OK 0
#define KBSRC_EINVAL 1   /* std::invalid_argument in the reference */
#define KBSRC_ERUNTIME 2 /* std::runtime_error: device errors, unsupported plans */
#define KBSRC_ENOMEM 3   /* std::bad_alloc: host or device memory exhausted */

/* Precision values, as kbsrc::Precision (include/kbsrc/kbsrc.hpp:8-12). LongDouble is rejected. */
#define KBSRC_FAST 0
#define KBSRC_EXACT 1
#define KBSRC_LONG_DOUBLE 2

/* kbsrc::Options (include/kbsrc/kbsrc.hpp:14-21) */
typedef struct kbsrc_options {
    int precision;  /* KBSRC_EXACT (default) or KBSRC_FAST */
    double length;  /* one-sided kernel length, seconds (default 0.2) */
    double beta;    /* Kaiser beta (default 40) */
    int threads;    /* accepted as the reference's worker count (>= 0); the GPU engine does not use it */
    int block;      /* block size factor nb, 0 = automatic */
    int lpc_edges;  /* Burg LPC edge extension, default 1 */
} kbsrc_options;

/* kbsrc::Stats (include/kbsrc/kbsrc.hpp:23-31) */
typedef struct kbsrc_stats {
    long long p, q, taps;
    int block, threads, general_fold;
    double design_seconds, process_seconds, passband, stopband;
    const char* isa;  /* "gfx950" */
} kbsrc_stats;

/* Fills the reference's defaults (include/kbsrc/kbsrc.hpp:14-21). */
void kbsrc_default_options(kbsrc_options* opt);

/* kbsrc::version (include/kbsrc/kbsrc.hpp:45) */
const char* kbsrc_version(void);

/* Message of the last failed call on this thread. */
const char* kbsrc_last_error(void);

/* Output frames per channel, ceil(frames * P / Q) (include/kbsrc/kbsrc.hpp:33-34, src/resampler.cpp:100-105). */
int kbsrc_output_frames(long long frames, int rate_in, int rate_out, long long* nout);

/* kbsrc::resample (include/kbsrc/kbsrc.hpp:41-43; src/dispatch.cpp:70-81): interleaved host input (frames x
 * channels) -> planar host output out[c * nout + i], nout = kbsrc_output_frames(). Exact runs the host-side dd
 * work in round-to-nearest and restores the caller's rounding mode. stats may be NULL. */
int kbsrc_resample(const double* interleaved, long long frames, int channels, int rate_in, int rate_out,
                   const kbsrc_options* opt, double* out, kbsrc_stats* stats);

/* ---- The engine seam: the reference's resample_impl split at run_engine (src/resampler.cpp:92-220,
 * run_engine :34-88). A plan owns the kernel spectrum, twiddle tables and work buffers on the current device. */
typedef struct kbsrc_plan kbsrc_plan;

/* Plans the blocks, designs the kernel (taps and spectrum on the GPU) and builds the FFT tables
 * (src/resampler.cpp:92-167,187-206). */
int kbsrc_plan_create(long long frames, int channels, int rate_in, int rate_out, const kbsrc_options* opt,
                      kbsrc_plan** plan);

/* Plan integers, in this order: P, Q, nout, nb, marg, Nb, Mb, hop, nblk, o, tot, base, pre, E, L, Lk, c2, gen,
 * up, equal, tasks, batch, xstride (23 values). xstride is the channel stride of the padded device input. */
#define KBSRC_PLAN_FIELDS 23
int kbsrc_plan_query(const kbsrc_plan* plan, long long* values);

/* Staging (src/resampler.cpp:168-186): host interleaved input -> device padded planar buffer d_padded of
 * channels x xstride doubles (channel c at d_padded + c * xstride), with the Burg edge extensions. stream: a hipStream_t, or NULL for the default. */
int kbsrc_plan_stage(kbsrc_plan* plan, const double* interleaved, double* d_padded, void* stream);

/* Staging from interleaved input already in device memory (frames x channels doubles): the same result as
 * kbsrc_plan_stage. Only the edge frames the Burg recursion trains on are read back to the host. Lets a caller
 * overlap the input upload with kbsrc_plan_create, whose design runs on a non-blocking stream of its own. */
int kbsrc_plan_stage_device(kbsrc_plan* plan, const double* d_interleaved, double* d_padded, void* stream);

/* kbsrc::Stats of the plan (include/kbsrc/kbsrc.hpp:23-31), process_seconds 0. */
int kbsrc_plan_stats(const kbsrc_plan* plan, kbsrc_stats* stats);

/* The engine (src/resampler.cpp:34-88): every (channel, block) task, device to device. d_padded as staged;
 * d_out: channels x nout doubles, planar. Asynchronous on stream. */
int kbsrc_plan_execute(kbsrc_plan* plan, const double* d_padded, double* d_out, void* stream);

/* The same execute with no device allocation (no counterpart in the reference; for callers that cannot allocate:
 * code under HIP graph capture, tight serving loops). The work buffers come from one static device array of
 * KBSRC_NOALLOC_SCRATCH_BYTES per GPU, held from the first such call on that GPU until the process ends; tasks run
 * in as many launches as that array needs, with the bits of kbsrc_plan_execute. A plan whose work buffers for one
 * task exceed the array is rejected (KBSRC_EINVAL). Every call on a GPU shares that GPU's array: calls that overlap
 * in time on one GPU (other streams, other threads) race and must be serialised by the caller. Plans with equal
 * rates have no engine and are rejected as by kbsrc_plan_execute. Asynchronous on stream. */
#define KBSRC_NOALLOC_SCRATCH_BYTES (128LL << 20)
int kbsrc_plan_execute_noalloc(kbsrc_plan* plan, const double* d_padded, double* d_out, void* stream);

/* Same, synchronous, with HIP-event times in milliseconds of the five kernel passes summed over the launch:
 * ms[0] forward column, ms[1] forward row, ms[2] split/multiply/fold, ms[3] inverse column, ms[4] inverse row. */
int kbsrc_plan_execute_timed(kbsrc_plan* plan, const double* d_padded, double* d_out, void* stream, double* ms);

/* The split-precision form of the same conversion (no counterpart in the reference; not bit-exact): exact-mode plans
 * only; every sub-FFT whose first two Stockham stages are radix 4 runs them as one radix-16 stage on the int8
 * matrix cores with `slices` base-128 slices per operand (4-8, 10, 12, 14, 15, 16), the rest in dd. Asynchronous on
 * stream; the first call per slice count makes the stage tables (synchronous). */
int kbsrc_plan_execute_split(kbsrc_plan* plan, const double* d_padded, double* d_out, int slices, void* stream);

/* The same conversion with every output sample written as its double-double value before the final rounding to
 * double: d_out_pairs holds channels x nout (hi, lo) pairs, 2 nout doubles per channel, hi + lo rounding to the
 * sample kbsrc_plan_execute (slices 0) or kbsrc_plan_execute_split (slices > 0) writes. Exact-mode plans only. No
 * counterpart in the reference: it exposes the internal dd result (src/resampler.cpp:80-84 rounds it away) so that
 * its error can be stated against a high-precision restatement of the same blocks. Asynchronous on stream. */
int kbsrc_plan_execute_dd(kbsrc_plan* plan, const double* d_padded, double* d_out_pairs, int slices, void* stream);

void kbsrc_plan_destroy(kbsrc_plan* plan);

/* ---- Batches: many jobs (files, channel counts, rate pairs, lengths) in one device input buffer and one device
 * output buffer, with one set of options. Each job's output is bit-identical to its own kbsrc_resample. Jobs whose
 * plans share a shape run as one launch per pass over all their (job, channel, block) tasks; no counterpart in the
 * reference (its callers loop over files). */
typedef struct kbsrc_job {
    long long frames;
    int channels, rate_in, rate_out;
} kbsrc_job;
typedef struct kbsrc_batch kbsrc_batch;

/* Plans every job (the reference's rejections, src/resampler.cpp:94-155) and designs one kernel per shape. */
int kbsrc_batch_create(int njobs, const kbsrc_job* jobs, const kbsrc_options* opt, kbsrc_batch** batch);

/* job >= 0: in_offset, in_doubles, out_offset, out_doubles, nout, group (-1 for an equal-rate copy); offsets in
 * doubles into the batch's device buffers; job's output is planar, channel c at out_offset + c * nout.
 * job == -1: fused, input buffer doubles, 0, output buffer doubles, jobs, groups; fused = 1 when executes run one
 * launch per step over every group (asked for with KBSRC_BATCH_FUSE=1 when the batch is created, and only when
 * every group fits the generic pass kernel). */
#define KBSRC_BATCH_FIELDS 6
int kbsrc_batch_query(const kbsrc_batch* batch, int job, long long* values);

/* Stages job's interleaved input (host, or device memory when on_device) into the batch input buffer d_in. */
int kbsrc_batch_stage(kbsrc_batch* batch, int job, const double* interleaved, int on_device, double* d_in,
                      void* stream);

/* Every job, device to device, asynchronous on stream (the groups fan out over the batch's own streams). */
int kbsrc_batch_execute(kbsrc_batch* batch, const double* d_in, double* d_out, void* stream);

void kbsrc_batch_destroy(kbsrc_batch* batch);

/* ---- Captured executes (no counterpart in the reference: its engine is host code). One plan or batch execute on
 * fixed device buffers, captured once into an instantiated HIP graph and replayed with one launch: for serving
 * loops that convert into the same buffers again and again. The graph uses the plan's (batch's) work buffers, so
 * it must not run concurrently with another execute or replay of the same plan, and the plan must outlive it.
 * Replays give the bits of kbsrc_plan_execute / kbsrc_batch_execute. */
typedef struct kbsrc_graph kbsrc_graph;
int kbsrc_plan_capture(kbsrc_plan* plan, const double* d_padded, double* d_out, kbsrc_graph** graph);
int kbsrc_batch_capture(kbsrc_batch* batch, const double* d_in, double* d_out, kbsrc_graph** graph);
int kbsrc_graph_launch(kbsrc_graph* graph, void* stream);   /* asynchronous on stream */
void kbsrc_graph_destroy(kbsrc_graph* graph);

/* ---- Measured kernel shapes (no counterpart in the reference). Runs the plan or batch once, as its execute call
 * does (same outputs), timing for every sub-FFT pass not measured before in this process its candidate shapes
 * (registered specialised variants; generic group sizes and thread counts) and keeping the fastest for every later
 * execute in the process. Any shape computes the same bits. Synchronous. */
int kbsrc_plan_tune(kbsrc_plan* plan, const double* d_padded, double* d_out, void* stream);
int kbsrc_batch_tune(kbsrc_batch* batch, const double* d_in, double* d_out, void* stream);
int kbsrc_tuned_shapes(void);   /* pass shapes measured so far in this process */

/* ---- Split-precision (Ozaki) radix-16 DFT on the int8 matrix cores, beside the plain dd pass (no counterpart in the
 * reference; the bit-exact conversion computes in dd; kbsrc_plan_execute_split uses this stage). nvec vectors of 16
 * complex dd values in device memory, element format {re.hi, im.hi, re.lo, im.lo}; y = forward DFT of each.
 * slices: 4-8, 10, 12, 14, 15, 16 base-128 digits per operand. Synchronous; the tables are made on the first call
 * per device (and slice count) and kept for the process. */
int kbsrc_split_dft16(const double* d_x, long long nvec, int slices, double* d_y, void* stream);
int kbsrc_plain_dft16(const double* d_x, long long nvec, double* d_y, void* stream);
/* The plain dd stage with the exchange between its two radix-4 stages done inside the wave (v_permlane32_swap)
 * instead of through LDS (no counterpart in the reference: a measured alternative of the plain pass's LDS round
 * trip; same bits as kbsrc_plain_dft16). Synchronous. */
int kbsrc_plain_dft16_inwave(const double* d_x, long long nvec, double* d_y, void* stream);

/* ---- The kernel design on the GPU (host buffers in and out; the plan calls do the same device to device). */

/* kernel_taps (src/design.hpp:13-15, src/design.cpp:17-37): one-sided Kaiser-sinc taps h[0..L], relative cutoff
 * a/b (L >= 1, a, b >= 1, beta > 0), as (hi, lo) pairs: out[2n] = h[n].hi, out[2n + 1] = h[n].lo, 2 (L + 1)
 * doubles. */
int kbsrc_kernel_taps(int L, long long a, long long b, double beta, double* out);

/* kernel_spectrum (src/design.hpp:17-18, src/design.cpp:40-49): taps h[0..Lk] as (hi, lo) pairs, placed
 * circularly in even length n > Lk; out: the n/2 + 1 values Re X[k] / scale as (hi, lo) pairs. */
int kbsrc_kernel_spectrum(const double* taps, int Lk, int n, double scale, double* out);

/* ---- Output packing on the GPU, for streamed WAV writing (src/io/wav.cpp:52-86,235-250, without dither):
 * frames [frame0, frame0 + frames) of the planar device output (channels x nout doubles) -> interleaved
 * little-endian sample bytes at d_bytes (device, 8-byte aligned), frames x channels x bytes-per-sample. Integer
 * formats round to nearest even, clip and add the number of clipped samples to *d_clipped (device); NaN -> 0.
 * Asynchronous on stream. */
#define KBSRC_PACK_F64 0
#define KBSRC_PACK_F32 1
#define KBSRC_PACK_S32 2
#define KBSRC_PACK_S24 3
#define KBSRC_PACK_S16 4
int kbsrc_pack(const double* d_planar, int channels, long long nout, long long frame0, long long frames, int format,
               void* d_bytes, unsigned long long* d_clipped, void* stream);

#ifdef __cplusplus
}
#endif
#endif
