// Host-side planning of the resampler: block plan, twiddle tables, Burg extension.
// Everything here is deterministic IEEE FP64 / double-double work that the reference also does on the host.
#pragma once
#include <string>
#include <vector>

#include "dd.hpp"
#include "design_math.hpp"   // cis, dd_sqrt, bessel_i0, kaiser_tap (host and device)

namespace kbsrc::gpu {

// The block plan (src/resampler.cpp:94-167,187-209). Integer for integer equal to the reference's, because the
// block size nb changes the rounding of the output.
struct Plan {
    long long fi = 0, fo = 0, P = 0, Q = 0, nin = 0, nout = 0;
    int ch = 0;
    bool equal = false, up = false, gen = false, fast = false;
    long long fk = 0, mn = 0, Lk = 0, c2 = 0, gc = 0, L = 0;
    double fml = 0;
    int nb = 0, marg = 0, Nb = 0, Mb = 0, hop = 0;
    long long nblk = 0, base = 0, tot = 0, E = 0, pre = 0, span = 0, o = 0;
    long long xstride = 0;   // doubles between channels of the padded device input: tot rounded up to 256 bytes
};

// Throws std::invalid_argument with the reference's messages.
Plan make_plan(long long frames, int channels, int rate_in, int rate_out, bool fast, double length, double beta,
               int block);

bool smooth7(long long n);
long long next_smooth7(long long n);

// cis(2 pi k / N) from two short tables and one dd complex product (src/dsp/fft.hpp:29-43)
class CisTable {
public:
    explicit CisTable(long long n);
    void get(long long k, dd& c, dd& s) const;
    long long block() const { return B_; }
    // the two short tables as {c.hi, c.lo, s.hi, s.lo} per entry: a[j] = cis(jB), b[l] = cis(l)

private:
    long long N_ = 0, B_ = 1;
    std::vector<dd> ac_, as_, bc_, bs_;
};

// Burg LPC prediction of n samples after x[0..len) (src/lpc.cpp:13-41): order 32 trained on the last 4096
// samples; zeros if the segment is silent, too short or the prediction unstable.
std::vector<double> burg_extrap(const double* x, long long len, long long n, int order = 32, int train = 4096);

// The 32-sample raised-cosine fade applied to the end of each extension (src/resampler.cpp:170-171)
void edge_fade(double fd[32]);

// Head and tail edge extensions of one channel, each E + 32 samples, faded (src/resampler.cpp:176-185).
// head[i] is the sample at position -1-i before the signal; tail[i] the sample at nin + i.
void edge_extensions(const Plan& p, const double* interleaved, int channel, std::vector<double>& head,
                     std::vector<double>& tail);
// The same from the first and the last edge_frames(p) interleaved frames only.
void edge_extensions(const Plan& p, const double* first, const double* last, int channel, std::vector<double>& head,
                     std::vector<double>& tail);
// Frames each edge's prediction trains on: min(frames, 4096) (src/lpc.cpp:13-41, src/resampler.cpp:176-185).
inline long long edge_frames(const Plan& p) { return p.nin < 4096 ? p.nin : 4096; }
// One of them: the tail (at_tail) or the head extension of one channel.
std::vector<double> edge_extension(const Plan& p, const double* first, const double* last, int channel, bool at_tail);

}  // namespace kbsrc::gpu
