"""Benchmark: exact-mode (double-double) output Msamples/s of the overlap-save resampler on one MI355X.

Workload: stereo 96000 Hz -> 44100 Hz, 180 s (17,280,000 frames per channel) of float64 uniform noise in
[-0.5, 0.5] (channel c from numpy default_rng(c + 1)), Kaiser beta 40, half-length 0.2 s, LPC edges, the
reference's automatic block (nb = 630, 214 blocks). One step = the engine over every (channel, block) task, from
the padded input resident in HBM to the output in HBM (the reference's `process_seconds` region).

Prints one JSON line. Beside the GPU figure it times the reference's own CPU engine (oracle/_ref, built from the
reference's sources) on the same workload on this host, and reports the roofline of the dominant kernel.
"""
import argparse
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

RATE_IN, RATE_OUT, SECONDS, CH = 96000, 44100, 180, 2
INT8_PEAK_OPS = 5.0e15   # MI355X dense int8 matrix peak
FP64_PEAK_TFLOPS = 78.6   # MI355X published FP64 peak
SPLIT_SLICES_2M100 = 15   # fewest base-128 slices whose radix-16 stage error is within 2^-100 (test_gpu_split)
WARM_MIN = 30   # fewest untimed executes before a timed region
HBM_PEAK_GBS = 8000.0
PEAK_CLOCK_MHZ = 2400.0   # MI355X peak engine clock, the clock the published peaks are quoted at


# ---- algorithmic FP64 flop counts of the reference's plain-dd / FP64 algorithm --------------------------------
def _radices(n):
    rs, t = [], n
    while t % 4 == 0:
        rs.append(4); t //= 4
    while t % 2 == 0:
        rs.append(2); t //= 2
    for f in (3, 5, 7):
        while t % f == 0:
            rs.append(f); t //= f
    assert t == 1
    return rs


def _costs(dd, accurate):
    """(add, mul) flop costs: dd mul 10 (FMA = 2), dd add 11 (one-TwoSum) or 20 (accurate); FP64 1 and 1."""
    if not dd:
        return 1, 1
    return (20 if accurate else 11), 10


def sfft_flops(n, dd, accurate):
    add, mul = _costs(dd, accurate)
    cmul = 4 * mul + 2 * add
    total, rem, s = 0, n, 1
    for r in _radices(n):
        m = rem // r
        nb, tw = m * s, (m - 1) * s   # butterflies, twiddled butterflies (p != 0)
        if r == 4:
            total += nb * 16 * add + tw * 3 * cmul
        elif r == 2:
            total += nb * 4 * add + tw * cmul
        else:
            h = (r - 1) // 2
            total += nb * ((6 * h + 4 * h * h + 4 * h) * add + 4 * h * h * mul) + tw * (r - 1) * cmul
        rem //= r
        s *= r
    return total


def cfft_pass_flops(n, dd):
    """Flops of the (column, row) passes of the reference's complex FFT of length n (four-step above 2048)."""
    if n > 2048:
        n1 = max(d for d in range(1, int(n ** 0.5) + 1) if n % d == 0)
        n2 = n // n1
        if n1 >= 8:
            add, mul = _costs(dd, False)
            col = n2 * sfft_flops(n1, dd, False) + (n1 - 1) * n2 * (4 * mul + 2 * add)
            row = n1 * sfft_flops(n2, dd, False)
            return col, row
    return 0, sfft_flops(n, dd, True)


def block_flops(Nb, Mb, dd, up):
    """Per (channel, block) task: forward col/row, split+multiply+fold, inverse col/row."""
    add, mul = _costs(dd, True)
    fc, fr = cfft_pass_flops(Nb // 2, dd)
    ic, ir = cfft_pass_flops(Mb // 2, dd)
    K = Mb // 2 + 1
    fold = 0 if up else max(0, min(Nb // 2, Mb - 1) - K + 1)
    mid = (Nb // 4) * (12 * add + 12 * mul) + (Mb // 4) * (12 * add + 8 * mul) + K * 2 * mul + fold * (2 * mul + 2 * add)
    return {"fwd_col": fc, "fwd_row": fr, "mid": mid, "inv_col": ic, "inv_row": ir + Mb}


def pass_flops(fl, p):
    """Flops of pass p from block_flops' five; row_mid, the forward row pass fused with the split/fold, does both."""
    return fl["fwd_row"] + fl["mid"] if p == "row_mid" else fl[p]


def block_bytes(Nb, Mb, dd, hop_out):
    """Compulsory HBM bytes per task of each pass: every pass reads its input and writes its output once. Complex
    elements are 32 B (dd) or 16 B; the forward pass reads Nb real float64 samples, the last pass writes the hop's
    hop_out float64 samples. The kernel spectrum H read by the split/multiply/fold is shared by all tasks and is
    added once per launch by the caller. row_mid: the forward row pass fused with the split/multiply/fold (it reads
    the column pass's output and writes the inverse FFT's input; the forward spectrum stays in LDS)."""
    elt = 32 if dd else 16
    h, hi = Nb // 2, Mb // 2
    return {"fwd_col": 8 * Nb + elt * h, "fwd_row": 2 * elt * h, "mid": elt * (h + hi), "row_mid": elt * (h + hi),
            "inv_col": 2 * elt * hi, "inv_row": elt * hi + 8 * hop_out}


# ---- rocprofv3 summaries ---------------------------------------------------------------------------------------
PASSES = ("fwd_col", "fwd_row", "mid", "row_mid", "inv_col", "inv_row")


def _csv_rows(path):
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _short_kernel(name):
    name = name.replace("kbsrc::gpu::", "").replace("(anonymous namespace)::", "").replace("void ", "")
    return re.sub(r"\(.*\)$", "", name).replace("Radices<", "R<").replace(", ", ",")


def _pass_of_dispatches(names):
    """Engine dispatch order per execute (no Bluestein) is fwd_col, fwd_row, mid, inv_col, inv_row, or fwd_col,
    row_mid, inv_col, inv_row where the forward row pass and the split/fold run as one kernel (k_rowmid): label each
    dispatch by its offset from the nearest k_mid or k_rowmid, where the others are the plain form's kernels
    (executes of the split form, whose forward rows run another kernel, are left out)."""
    lab = [None] * len(names)

    def plain_pass(n):
        return "k_fixed<" in n and "k_fixed_split" not in n

    for i, n in enumerate(names):
        if "k_mid<" in n and i >= 2 and i + 2 < len(names) and all(plain_pass(names[i + o]) for o in (-2, -1, 1, 2)):
            for off, p in zip((-2, -1, 0, 1, 2), ("fwd_col", "fwd_row", "mid", "inv_col", "inv_row")):
                lab[i + off] = p
        if "k_rowmid<" in n and i >= 1 and i + 2 < len(names) and all(plain_pass(names[i + o]) for o in (-1, 1, 2)):
            for off, p in zip((-1, 0, 1, 2), ("fwd_col", "row_mid", "inv_col", "inv_row")):
                lab[i + off] = p
    return lab


def summarize_rocprof(d, mode, steps=20, warmup=5):
    """Reads <d>/stats (--kernel-trace --stats), <d>/fetch (--pmc FETCH_SIZE) and <d>/write (--pmc WRITE_SIZE),
    all csv, and prints the per-pass summary kept under profiles/ (HBM bytes scaled by what each counter reads for
    the 256 MiB copy of --profile-copy)."""
    trace = _csv_rows(os.path.join(d, "stats", "run_kernel_trace.csv"))
    names = [r["Kernel_Name"] for r in trace]
    lab = _pass_of_dispatches(names)
    dur, kname = {p: [] for p in PASSES}, {}
    for r, p in zip(trace, lab):
        if p:
            dur[p].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
            kname[p] = _short_kernel(r["Kernel_Name"])
    # the timed region of the profiled command: plain executes max(W, WARM_MIN) .. + K - 1, in dispatch order
    t0 = max(warmup, WARM_MIN)
    timed = {p: v[t0:t0 + steps] for p, v in dur.items()}
    copy_bytes = (1 << 25) * 8
    cnt, factor = {}, {}
    for c, sub in (("FETCH_SIZE", "fetch"), ("WRITE_SIZE", "write")):
        rows = _csv_rows(os.path.join(d, sub, "run_counter_collection.csv"))
        rows = [r for r in rows if r["Counter_Name"] == c]
        copies = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]]
        factor[c] = copy_bytes / 1024 / max(copies)
        lab_c = _pass_of_dispatches([r["Kernel_Name"] for r in rows])
        cnt[c] = {p: [float(r["Counter_Value"]) for r, q in zip(rows, lab_c) if q == p] for p in PASSES}
    out = []
    out.append(f"# Engine passes, stereo 96000 -> 44100 Hz, 180 s, {mode} mode: MI355X, rocprofv3 (csv output)")
    out.append("# commands, from /tmp with TMPDIR=/tmp, R = the repository root, D = an output folder:")
    out.append(f"#   rocprofv3 --kernel-trace --stats -d D/stats -o run --output-format csv -- python3 R/bench.py "
               f"--mode {mode} --steps {steps} --warmup {warmup} --no-cpu")
    for c, sub in (("FETCH_SIZE", "fetch"), ("WRITE_SIZE", "write")):
        out.append(f"#   rocprofv3 --pmc {c} -d D/{sub} -o run --output-format csv -- python3 R/bench.py --mode {mode} "
                   "--steps 2 --warmup 1 --no-cpu --no-split --cpu-runs 1 --profile-copy")
    out.append(f"#   python3 R/bench.py --summarize-rocprof D --mode {mode} --steps {steps} --warmup {warmup}")
    out.append(f"# calibration, 256 MiB device copy: FETCH_SIZE x {factor['FETCH_SIZE']:.3f}, "
               f"WRITE_SIZE x {factor['WRITE_SIZE']:.3f} (counter KiB -> bytes moved)")
    out.append(f"# avg_us: the mean over the {steps} executes of the timed region (executes {t0}..{t0 + steps - 1} of "
               f"the run, after {t0} untimed ones); all_avg_us: over every execute of the run (the warm-up and the "
               f"executes after host-side pauses included)")
    out.append(f"# {'pass':8s} {'launches':>8s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'fetch_MB':>9s} "
               f"{'write_MB':>9s} {'GB/s':>8s} {'all_avg_us':>10s}  kernel")
    for p in PASSES:
        t, ta = timed[p], dur[p]
        if not t:
            continue
        fb = sum(cnt["FETCH_SIZE"][p]) / len(cnt["FETCH_SIZE"][p]) * 1024 * factor["FETCH_SIZE"]
        wb = sum(cnt["WRITE_SIZE"][p]) / len(cnt["WRITE_SIZE"][p]) * 1024 * factor["WRITE_SIZE"]
        avg = sum(t) / len(t)
        out.append(f"  {p:8s} {len(t):8d} {avg:9.1f} {min(t):9.1f} {max(t):9.1f} {fb / 1e6:9.1f} {wb / 1e6:9.1f} "
                   f"{(fb + wb) / avg / 1e3:8.0f} {sum(ta) / len(ta):10.1f}  {kname[p]}")
    out.append(f"timed_sum_us {sum(sum(timed[p]) / len(timed[p]) for p in PASSES if timed[p]):.1f}")
    same = os.path.join(d, "bench.json")   # the unprofiled bench line of the same box, if the run kept one
    if os.path.exists(same):
        with open(same) as f:
            b = json.loads(f.read().strip().splitlines()[-1])
        ev = " ".join(f"{k} {v['ms'] * 1e3:.1f}" for k, v in b["passes"].items())
        ul = b["roofline"].get("under_load")
        out.append(f"# the same box without the profiler (python3 R/bench.py --mode {mode} --no-cpu > D/bench.json): "
                   f"{b['ms_per_step']} ms/step; pass events (us): {ev}" +
                   (f"; under load {ul['socket_power_w']:.0f} W, GFX clock {ul['gfx_clock_mhz']:.0f} MHz" if ul else ""))
    for p in PASSES:
        if cnt["FETCH_SIZE"][p]:
            fb = sum(cnt["FETCH_SIZE"][p]) / len(cnt["FETCH_SIZE"][p]) * 1024 * factor["FETCH_SIZE"]
            wb = sum(cnt["WRITE_SIZE"][p]) / len(cnt["WRITE_SIZE"][p]) * 1024 * factor["WRITE_SIZE"]
            out.append(f"traffic {p} {int(fb + wb)}")
    print("\n".join(out))
    return 0


def summarize_split(d):
    """Reads <d> (rocprofv3 --kernel-trace --stats of the split workload, csv) and prints the radix-16 stage
    summary kept under profiles/: per kernel, launches and durations, and the per-launch times of the conversion's
    320-point forward row passes, plain and split."""
    rows = _csv_rows(os.path.join(d, "run_kernel_trace.csv"))
    dur = {}
    for r in rows:
        dur.setdefault(r["Kernel_Name"], []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    gb = (1 << 22) * 16 * 64 / 1e9   # 32 B in + 32 B out per point (one dd complex each way)
    stage, conv = [], []
    for name, v in dur.items():
        m = re.search(r"k_split_dft16<(\d+)>", name)
        if m:
            stage.append((2, int(m.group(1)), f"split radix-16, S = {m.group(1)}", v))
        elif "k_dft16_inwave" in name:
            stage.append((1, 0, "plain dd radix-16, in-wave exchange", v))
        elif "Radices<4, 4>," in name:
            stage.append((0, 0, "plain dd radix-16 pass (2 x radix 4, LDS)", v))
        elif "Radices<4, 4, 4, 5>" in name and ("k_fixed_split" in name or "k_fixed<" in name):
            m = re.search(r"k_fixed_split<.*, (\d+)>\(", name)
            conv.append((int(m.group(1)) if m else 0,
                         f"split forward row pass, S = {m.group(1)}" if m else "plain forward row pass (320 points)", v))
        elif "Radices<4, 4, 4, 5>" in name and "k_rowmid<" in name:
            conv.append((-1, "plain forward row pass + split/fold (one kernel)", v))
    out = ["# Split-precision radix-16 DFT (int8 MFMA slices), the plain dd radix-16 pass and the plain pass with its "
           "exchange inside the wave:",
           "# 4,194,304 vectors of 16 complex dd values (67,108,864 points, 32 B in + 32 B out per point = 4.295 GB "
           "per launch); MI355X, rocprofv3 (csv output).",
           "# commands, from /tmp with TMPDIR=/tmp, R = the repository root, D = an output folder:",
           "#   rocprofv3 --kernel-trace --stats -d D -o run --output-format csv -- python3 R/bench.py --workload split "
           "--steps 3 --warmup 1 --no-cpu",
           "#   python3 R/bench.py --summarize-split D",
           "# (the same run also times the 180 s 96 kHz -> 44.1 kHz conversion, plain and in the split form; its row "
           "passes are listed last)",
           f"# {'kernel':44s} {'launches':>8s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'GB/s(algorithmic)':>18s}"]
    for _, _, label, v in sorted(stage):
        avg = sum(v) / len(v)
        out.append(f"  {label:44s} {len(v):8d} {avg:9.1f} {min(v):9.1f} {max(v):9.1f} {gb / (avg / 1e6):18.0f}")
    out.append("# conversion row passes (per launch; 214 tasks, sub-FFTs of 320 points; the plain form runs its forward rows "
               "and the split/fold step as one kernel, the split form its split row pass and then the split/fold kernel):")
    for _, label, v in sorted(conv):
        out.append(f"  {label:44s} {len(v):8d} {sum(v) / len(v):9.1f} {min(v):9.1f} {max(v):9.1f}")
    print("\n".join(out))
    return 0


COUNTER_SETS = (
    ("pmc1", "SQ_WAVES SQ_INSTS_VALU SQ_INSTS_VALU_ADD_F64 SQ_INSTS_VALU_MUL_F64 SQ_INSTS_VALU_FMA_F64 SQ_INSTS_LDS "
             "SQ_INSTS_SALU SQ_ACTIVE_INST_VALU GRBM_GUI_ACTIVE"),
    ("pmc2", "SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_WAIT_INST_LDS SQ_WAIT_INST_ANY SQ_ACTIVE_INST_LDS SQ_LDS_BANK_CONFLICT "
             "SQ_ACTIVE_INST_ANY SQ_THREAD_CYCLES_VALU GRBM_GUI_ACTIVE"),
)


def summarize_kernels(d, top=30):
    """Reads <d>/run_kernel_stats.csv (rocprofv3 --kernel-trace --stats, csv) and prints the kernels by total time:
    calls, average, minimum and maximum duration, total (the per-kernel summary kept under profiles/ for workloads
    without a per-pass summary)."""
    rows = _csv_rows(os.path.join(d, "run_kernel_stats.csv"))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = [f"# {'kernel':70s} {'calls':>6s} {'avg_us':>10s} {'min_us':>10s} {'max_us':>10s} {'total_ms':>10s}"]
    for r in rows[:top]:
        out.append(f"  {_short_kernel(r['Name'])[:70]:70s} {int(r['Calls']):6d} {float(r['AverageNs']) / 1e3:10.1f} "
                   f"{float(r['MinNs']) / 1e3:10.1f} {float(r['MaxNs']) / 1e3:10.1f} "
                   f"{float(r['TotalDurationNs']) / 1e6:10.2f}")
    print("\n".join(out))
    return 0


def summarize_counters(d, mode):
    """Reads <d>/pmc1 and <d>/pmc2 (rocprofv3 --pmc, csv) and prints the per-pass VALU and LDS summary kept under
    profiles/: counters summed over XCDs and shader engines per dispatch, averaged over the dispatches of a pass."""
    per = {}
    for sub, _ in COUNTER_SETS:
        rows = _csv_rows(os.path.join(d, sub, "run_counter_collection.csv"))
        disp = {}
        for r in rows:
            key = (int(r["Dispatch_Id"]), r["Kernel_Name"])
            disp.setdefault(key, {}).setdefault(r["Counter_Name"], 0.0)
            disp[key][r["Counter_Name"]] += float(r["Counter_Value"])
        keys = sorted(disp)
        lab = _pass_of_dispatches([k[1] for k in keys])
        for k, p in zip(keys, lab):
            if p:
                acc = per.setdefault(p, {})
                for c, v in disp[k].items():
                    acc.setdefault(c, []).append(v)
    avg = {p: {c: sum(v) / len(v) for c, v in cs.items()} for p, cs in per.items()}
    out = [f"# VALU and LDS counters per pass, stereo 96000 -> 44100 Hz, 180 s, {mode} (214 tasks per launch); MI355X, "
           "rocprofv3.",
           "# commands, from /tmp with TMPDIR=/tmp, R = the repository root, D = an output folder (two passes: at most "
           "8 SQ counters each):"]
    for sub, counters in COUNTER_SETS:
        out.append(f"#   rocprofv3 --pmc {counters} -d D/{sub} -o run --output-format csv -- python3 R/bench.py --mode "
                   f"{mode} --steps 2 --warmup 1 --no-cpu --no-split --cpu-runs 1")
    out.append(f"#   python3 R/bench.py --summarize-counters D --mode {mode}")
    out.append("# per dispatch, summed over all XCDs/SEs, averaged over dispatches. VALU busy = SQ_ACTIVE_INST_VALU x 4 "
               "cycles /")
    out.append("# (1024 SIMDs x GRBM_GUI_ACTIVE / 8 XCDs): SQ_ACTIVE_INST_VALU counts one quad-cycle per wave64 VALU "
               "instruction (FP64 at full rate on 16 lanes per SIMD).")
    out.append(f"# {'pass':8s} {'VALU insts':>10s} {'FP64 share':>10s} {'adds':>6s} {'muls':>6s} {'fmas':>6s} "
               f"{'VALU busy':>9s} {'LDS insts':>10s} {'LDS bank-conflict/LDS active':>28s}")
    for p in PASSES:
        a = avg.get(p)
        if not a:
            continue
        v = a["SQ_INSTS_VALU"]
        f64 = a["SQ_INSTS_VALU_ADD_F64"] + a["SQ_INSTS_VALU_MUL_F64"] + a["SQ_INSTS_VALU_FMA_F64"]
        busy = a["SQ_ACTIVE_INST_VALU"] * 4 / (1024 * a["GRBM_GUI_ACTIVE"] / 8)
        conf = a["SQ_LDS_BANK_CONFLICT"] / a["SQ_ACTIVE_INST_LDS"] if a.get("SQ_ACTIVE_INST_LDS") else 0.0
        out.append(f"  {p:8s} {v:10.4g} {f64 / v:10.2f} {a['SQ_INSTS_VALU_ADD_F64'] / v:6.2f} "
                   f"{a['SQ_INSTS_VALU_MUL_F64'] / v:6.2f} {a['SQ_INSTS_VALU_FMA_F64'] / v:6.2f} {busy:9.2f} "
                   f"{a['SQ_INSTS_LDS']:10.4g} {conf:28.2f}")
    out.append("")
    out.append(f"# {'pass':6s} {'wave_cycles':>11s} {'wait_any/wave':>13s} {'wait_lds/wave':>13s} {'active_any/wave':>15s} "
               "thread_cycles_valu/(64*active_valu)")
    for p in PASSES:
        a = avg.get(p)
        if not a:
            continue
        wc = a["SQ_WAVE_CYCLES"]
        out.append(f"  {p:8s} {wc:11.4g} {a['SQ_WAIT_INST_ANY'] / wc:13.3f} {a['SQ_WAIT_INST_LDS'] / wc:13.3f} "
                   f"{a['SQ_ACTIVE_INST_ANY'] / wc:15.3f} "
                   f"{a['SQ_THREAD_CYCLES_VALU'] / (64 * a['SQ_ACTIVE_INST_VALU']) if 'SQ_ACTIVE_INST_VALU' in a else 0:10.3f}")
    print("\n".join(out))
    return 0


def committed_traffic(mode, pass_name):
    """HBM bytes per launch of a pass from the committed rocprofv3 summary, or None."""
    path = os.path.join(ROOT, "profiles", f"passes_96k_to_44k1_180s_{mode}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic" and f[1] == pass_name:
            return int(f[2])
    return None


def committed_avg_us(mode, pass_name):
    """Average duration (us) of a pass's kernel over its launches in the committed rocprofv3 summary, or None."""
    path = os.path.join(ROOT, "profiles", f"passes_96k_to_44k1_180s_{mode}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 9 and f[0] == pass_name and f[1].isdigit():
            return float(f[2])
    return None


def committed_valu_busy(mode, pass_name):
    """Fraction of cycles a pass keeps the VALU issuing, from the committed rocprofv3 counter summary, or None."""
    path = os.path.join(ROOT, "profiles", f"valu_counters_96k_to_44k1_180s_{mode}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 9 and f[0] == pass_name:
            return float(f[6])
    return None


def committed_valu_insts(mode):
    """VALU wave-instructions of one execute (all passes) from the committed counter summary, or None."""
    path = os.path.join(ROOT, "profiles", f"valu_counters_96k_to_44k1_180s_{mode}.txt")
    if not os.path.exists(path):
        return None
    total = 0.0
    for line in open(path):
        f = line.split()
        if len(f) == 9 and f[0] in PASSES:
            total += float(f[1])
    return total or None


def host_cpu():
    """The host's CPU model and the logical CPUs this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def ref_engine_once(src, dst, threads, rate_out=RATE_OUT, extra=()):
    """One run of the reference's CLI (oracle/_ref/kbsrc, compiled from its sources): (resample s, design s) from
    its `-v` line, which times the engine without design and file I/O (app/main.cpp:144-145); None on failure."""
    ref = os.path.join(ROOT, "oracle", "_ref", "kbsrc")
    r = subprocess.run([ref, src, dst, "--rate", str(rate_out), "-t", str(threads), "-v", *extra],
                       capture_output=True, text=True, timeout=300)
    m = re.search(r"design ([0-9.]+)s\s+resample ([0-9.]+)s", r.stderr)
    if r.returncode or not m:
        return None
    return float(m.group(2)), float(m.group(1))


def median(v):
    v = sorted(v)
    n = len(v)
    return v[n // 2] if n % 2 else 0.5 * (v[n // 2 - 1] + v[n // 2])


# ---- batched form: many files, channel counts and rate pairs in one batch ------------------------------------------
BATCH_RATES = [44100, 48000, 88200, 96000, 176400, 192000]


def batch_jobs(n=64, seed=4):
    """64 jobs: lengths uniform in [0.5, 300] s, channels in {1, 2, 6, 8}, rate pairs from BATCH_RATES both ways
    (never equal); even jobs uniform noise in [-0.5, 0.5], odd jobs a sum of 8 sinusoids (amplitude 0.1, random
    frequencies below 0.45 x the lower rate, random phases), every channel its own draw."""
    import numpy as np
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(n):
        secs = float(rng.uniform(0.5, 300.0))
        ch = int(rng.choice([1, 2, 6, 8]))
        ri, ro = (int(v) for v in rng.choice(BATCH_RATES, 2, replace=False))
        jobs.append({"frames": int(round(secs * ri)), "channels": ch, "rate_in": ri, "rate_out": ro,
                     "kind": "noise" if j % 2 == 0 else "sines", "seed": seed * 1000 + j})
    return jobs


def batch_jobs_short(n=256, seed=9):
    """256 short jobs: lengths uniform in [0.5, 2] s, 1 or 2 channels, rate pairs from BATCH_RATES both ways, signals
    as in batch_jobs."""
    import numpy as np
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(n):
        secs = float(rng.uniform(0.5, 2.0))
        ch = int(rng.choice([1, 2]))
        ri, ro = (int(v) for v in rng.choice(BATCH_RATES, 2, replace=False))
        jobs.append({"frames": int(round(secs * ri)), "channels": ch, "rate_in": ri, "rate_out": ro,
                     "kind": "noise" if j % 2 == 0 else "sines", "seed": seed * 1000 + j})
    return jobs


def batch_signal(job, max_frames=None):
    """The job's interleaved (frames, channels) float64 signal, generated on the GPU (torch)."""
    import numpy as np
    import torch
    frames = job["frames"] if max_frames is None else min(job["frames"], max_frames)
    ch = job["channels"]
    if job["kind"] == "noise":
        g = torch.Generator(device="cuda")
        g.manual_seed(job["seed"])
        return torch.rand((frames, ch), dtype=torch.float64, device="cuda", generator=g) - 0.5
    rng = np.random.default_rng(job["seed"])
    fmax = 0.45 * min(job["rate_in"], job["rate_out"])
    n = torch.arange(frames, dtype=torch.float64, device="cuda")
    x = torch.zeros((frames, ch), dtype=torch.float64, device="cuda")
    for c in range(ch):
        for f, ph in zip(rng.uniform(0, fmax, 8), rng.uniform(0, 2 * np.pi, 8)):
            x[:, c] += 0.1 * torch.sin(n * (2 * np.pi * f / job["rate_in"]) + ph)
    return x


def graph_timing(args, owner, d_src, d_out):
    """Time replays of the execute captured as a HIP graph (Plan.capture / Batch.capture) into a second output
    buffer, and check that they give d_out's bits (d_out holds an eager execute's output)."""
    import torch
    alt = torch.empty_like(d_out)
    g = owner.capture(d_src.data_ptr(), alt.data_ptr())
    for _ in range(max(args.warmup, WARM_MIN)):
        g.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.launch()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    same = bool(torch.equal(alt, d_out))
    g.close()
    del alt
    return ms, same


def bench_batch(args):
    import numpy as np
    import torch

    import kbsrc_amd
    jobs = batch_jobs_short() if args.batch_set == "short" else batch_jobs()
    t0 = time.perf_counter()
    b = kbsrc_amd.Batch([(j["frames"], j["channels"], j["rate_in"], j["rate_out"]) for j in jobs], "exact")
    t_design = time.perf_counter() - t0
    d_in = torch.empty(b.in_doubles, dtype=torch.float64, device="cuda")
    d_out = torch.empty(b.out_doubles, dtype=torch.float64, device="cuda")
    t0 = time.perf_counter()
    for k, job in enumerate(jobs):
        x = batch_signal(job)
        torch.cuda.synchronize()
        b.stage_device(k, x.data_ptr(), d_in.data_ptr())
        del x
    torch.cuda.synchronize()
    t_stage = time.perf_counter() - t0
    in_samples = sum(j["frames"] * j["channels"] for j in jobs)
    out_samples = sum(L["nout"] * j["channels"] for j, L in zip(jobs, b.layout))
    print(f"batch: {len(jobs)} jobs, {b.groups} shapes, {in_samples / 1e9:.2f} G input samples, "
          f"{out_samples / 1e9:.2f} G output samples; design {t_design:.2f} s, generate+stage {t_stage:.2f} s",
          file=sys.stderr, flush=True)
    t_tune = 0.0
    if args.tune:
        t0 = time.perf_counter()
        b.tune(d_in.data_ptr(), d_out.data_ptr())
        t_tune = time.perf_counter() - t0
        print(f"tuned {kbsrc_amd.tuned_shapes()} pass shapes in {t_tune:.2f} s", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        b.execute(d_in.data_ptr(), d_out.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.execute(d_in.data_ptr(), d_out.data_ptr())
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    value = out_samples / (ms / 1e3) / 1e6
    print(f"batched: {ms:.2f} ms per batch -> {value:.1f} output Msamples/s", file=sys.stderr, flush=True)
    graph_ms, graph_same = graph_timing(args, b, d_in, d_out)
    print(f"batched, captured graph replay: {graph_ms:.2f} ms per batch; same output: {graph_same}", file=sys.stderr,
          flush=True)

    # the same jobs one plan at a time over the same staged input (a job's padded layout is its plan's)
    ref_out = torch.empty_like(d_out)
    single_ms = 0.0
    flops = 0.0
    for k, (job, L) in enumerate(zip(jobs, b.layout)):
        plan = kbsrc_amd.Plan(job["frames"], job["channels"], job["rate_in"], job["rate_out"], "exact")
        info = plan.info
        fl = block_flops(info["Nb"], info["Mb"], True, bool(info["up"]))
        flops += sum(fl.values()) * info["tasks"]
        src = d_in.data_ptr() + 8 * L["in_offset"]
        dst = ref_out.data_ptr() + 8 * L["out_offset"]
        plan.execute(src, dst)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.execute(src, dst)
        torch.cuda.synchronize()
        single_ms += (time.perf_counter() - t0) * 1e3
        plan.close()
    same = all(torch.equal(d_out[L["out_offset"]:L["out_offset"] + L["out_doubles"]],
                           ref_out[L["out_offset"]:L["out_offset"] + L["out_doubles"]]) for L in b.layout)
    print(f"one plan at a time: {single_ms:.2f} ms; outputs identical: {same}", file=sys.stderr, flush=True)
    del ref_out

    cpu = None
    ref = os.path.join(ROOT, "oracle", "_ref", "kbsrc")
    if not args.no_cpu and os.path.exists(ref):
        from tests.helpers import wav_f64
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        eng, outs, names = 0.0, 0, []
        with tempfile.TemporaryDirectory(prefix="kbsrc-bench-") as td:
            for job in jobs[:4]:
                x = batch_signal(job, max_frames=20 * job["rate_in"]).cpu().numpy()
                src, dst = os.path.join(td, "in.wav"), os.path.join(td, "out.wav")
                with open(src, "wb") as f:
                    f.write(wav_f64(x, job["channels"], job["rate_in"]))
                r = subprocess.run([ref, src, dst, "--rate", str(job["rate_out"]), "-t", str(threads), "-v"],
                                   capture_output=True, text=True, timeout=300)
                m = re.search(r"design ([0-9.]+)s\s+resample ([0-9.]+)s", r.stderr)
                if r.returncode or not m:
                    eng = 0.0
                    break
                eng += float(m.group(2))
                nout = kbsrc_amd.output_frames(x.shape[0], job["rate_in"], job["rate_out"])
                outs += nout * job["channels"]
                names.append(f"{job['rate_in']}->{job['rate_out']} {job['channels']}ch {x.shape[0]} frames")
        if eng > 0:
            cpu = {"value": round(outs / eng / 1e6, 2), "unit": "output Msamples/s", "cores": threads,
                   "kind": "reference",
                   "sample": "the first 4 jobs cut to 20 s each (" + "; ".join(names) +
                             f"): reference engine resample time {eng:.3f} s, design and file I/O excluded"}
    line = {
        "metric": "batched exact-mode output Msamples/s (" + ("256 short jobs" if args.batch_set == "short"
                                                               else "64 mixed jobs") + ")",
        "value": round(value, 2),
        "unit": "output Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": ("256 jobs, 0.5-2 s, 1/2 channels" if args.batch_set == "short" else
                                "64 jobs, 0.5-300 s, 1/2/6/8 channels") +
                               ", rate pairs among 44.1/48/88.2/96/176.4/192 kHz both ways, half noise half 8 "
                               "sinusoids, exact mode, len 0.2, beta 40, LPC on",
                   "jobs": len(jobs), "shapes": b.groups, "input_samples": in_samples, "output_samples": out_samples,
                   "parallelism": "single GPU"},
        "graph_ms_per_step": round(graph_ms, 3),
        "graph_outputs_identical": graph_same,
        "one_plan_at_a_time_ms": round(single_ms, 3),
        "batch_speedup": round(single_ms / ms, 3),
        "outputs_identical": bool(same),
        "algorithmic_gflop_per_step": round(flops / 1e9, 2),
        "step_tflops": round(flops / 1e9 / ms, 3),
        "design_s": round(t_design, 3),
        "tuned": bool(args.tune),
        "tune_s": round(t_tune, 3),
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu_engine"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)
    b.close()
    return 0


# ---- extreme-ratio plans: 10 MHz -> 48 kHz mono 30 s and 8-channel 384 kHz -> 44.1 kHz 600 s, --len 0.5 -------
def bench_extreme(args):
    import numpy as np
    import torch

    import kbsrc_amd
    from tests import helpers as H

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev).cuda_stream
    cases = []

    def one(name, x, rate_in, rate_out, split_slices=()):
        frames, ch = x.shape
        t0 = time.perf_counter()
        plan = kbsrc_amd.Plan(frames, ch, rate_in, rate_out, "exact", length=0.5)
        t_design = time.perf_counter() - t0
        info = plan.info
        xx = torch.empty(ch * info["xstride"], dtype=torch.float64, device=dev)
        out = torch.empty(ch * info["nout"], dtype=torch.float64, device=dev)
        plan.stage(x, xx.data_ptr(), stream)
        torch.cuda.synchronize()

        def timed(run):
            for _ in range(max(1, args.warmup)):
                run()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                run()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) * 1e3 / args.steps

        ms = timed(lambda: plan.execute(xx.data_ptr(), out.data_ptr(), stream))
        kt = plan.execute_timed(xx.data_ptr(), out.data_ptr(), stream)
        nout_total = ch * info["nout"]
        # the same conversion in fast mode (FP64; its own automatic block)
        fplan = kbsrc_amd.Plan(frames, ch, rate_in, rate_out, "fast", length=0.5)
        fxx = torch.empty(ch * fplan.info["xstride"], dtype=torch.float64, device=dev)
        fplan.stage(x, fxx.data_ptr(), stream)
        fast_ms = timed(lambda: fplan.execute(fxx.data_ptr(), out.data_ptr(), stream))
        fplan.close()
        del fxx
        plan.execute(xx.data_ptr(), out.data_ptr(), stream)   # the exact output again, for the split comparison
        fl = block_flops(info["Nb"], info["Mb"], True, bool(info["up"]))
        gflop = sum(fl.values()) * info["tasks"] / 1e9
        row = {"workload": name, "ms": round(ms, 3), "output_msamples_s": round(nout_total / (ms / 1e3) / 1e6, 1),
               "tflops": round(gflop / ms, 2), "nb": info["nb"], "Nb": info["Nb"], "Mb": info["Mb"],
               "tasks": info["tasks"], "design_s": round(t_design, 3),
               "pass_ms": {k: round(v, 3) for k, v in kt.items()}, "fast_ms": round(fast_ms, 3)}
        print(f"{name}: {ms:.3f} ms -> {row['output_msamples_s']} output Msamples/s, {row['tflops']} TFLOP/s "
              f"(design {t_design:.2f} s); fast mode {fast_ms:.3f} ms", file=sys.stderr, flush=True)
        if split_slices:
            plain = out.clone()
            peak = float(plain.abs().max().item())
            row["split"] = {}
            for sl in split_slices:
                sms = timed(lambda: plan.execute_split(xx.data_ptr(), out.data_ptr(), sl, stream))
                d = float((out - plain).abs().max().item()) / peak
                row["split"][sl] = {"ms": round(sms, 3), "speed_vs_plain": round(ms / sms, 3),
                                    "max_diff_vs_plain_over_peak": d}
                print(f"  split S={sl}: {sms:.3f} ms (x{ms / sms:.3f}), max diff vs plain {d:.2e} of peak",
                      file=sys.stderr, flush=True)
            del plain
        plan.close()
        del xx, out
        torch.cuda.empty_cache()
        cases.append(row)

    one("mono 10000000 -> 48000 Hz, 30 s, exact, --len 0.5", H.input_10m_30s().reshape(-1, 1), 10000000, 48000,
        split_slices=(8, 15))
    one("8 ch 384000 -> 44100 Hz, 600 s, exact, --len 0.5", H.input_384k(H.FRAMES_384K_FULL), 384000, 44100,
        split_slices=(8, 15))
    cpu = None
    ref = os.path.join(ROOT, "oracle", "_ref", "kbsrc")
    if not args.no_cpu and os.path.exists(ref):   # bounded sample: 2 s of the 384 kHz input at the same block
        from tests.golden.make_extreme_pins import write_wav
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        with tempfile.TemporaryDirectory(prefix="kbsrc-bench-") as td:
            src, dst = os.path.join(td, "in.wav"), os.path.join(td, "out.wav")
            write_wav(src, H.FRAMES_384K_SHORT, 8, 384000, H.input_384k_chunks(H.FRAMES_384K_SHORT))
            r = subprocess.run([ref, src, dst, "--rate", "44100", "--len", "0.5", "--block", "1536", "-t",
                                str(threads), "-v"], capture_output=True, text=True, timeout=300)
            m = re.search(r"design ([0-9.]+)s\s+resample ([0-9.]+)s", r.stderr)
            if not r.returncode and m:
                eng = float(m.group(2))
                cpu = {"value": round(8 * 88200 / eng / 1e6, 2), "unit": "output Msamples/s", "cores": threads,
                       "kind": "reference",
                       "sample": f"8 ch x 2 s of the 384 kHz input at block 1536 (the full plan's transform "
                                 f"shapes): reference engine resample {eng:.3f} s, design and file I/O excluded"}
    big = cases[1]
    line = {
        "metric": "exact-mode output Msamples/s, extreme-ratio plans (8 ch 384k->44.1k 600 s headline)",
        "value": big["output_msamples_s"], "unit": "output Msamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": big["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": big["workload"], "parallelism": "single GPU"},
        "cases": cases, "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu_engine"] = round(big["output_msamples_s"] / cpu["value"], 1)
    print(json.dumps(line), flush=True)
    return 0


# ---- split-precision form: radix-16 DFT stages, int8 MFMA slices against the plain dd pass ---------------------
def bench_split(args):
    import numpy as np
    import torch

    import kbsrc_amd
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from test_gpu_split import dd_vectors, errors
    nvec = 1 << 22   # 4,194,304 radix-16 DFTs = 67,108,864 points
    x = torch.from_numpy(dd_vectors(nvec, 5).reshape(-1)).cuda()
    y = torch.empty_like(x)
    ref = torch.empty_like(x)
    sample = 16

    def timed(fn, *a):
        for _ in range(max(1, args.warmup)):
            fn(x.data_ptr(), nvec, *a, y.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn(x.data_ptr(), nvec, *a, y.data_ptr())
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.steps

    xs = x[: sample * 64].cpu().numpy().reshape(sample, 16, 4)
    plain_ms = timed(kbsrc_amd.plain_dft16)
    ref.copy_(y)
    inwave_ms = timed(kbsrc_amd.plain_dft16_inwave)
    inwave_same = bool(torch.equal(y, ref))
    plain_ms = min(plain_ms, timed(kbsrc_amd.plain_dft16))   # again, after the other: clocks as warm for both
    print(f"plain dd radix-16: LDS exchange {plain_ms:.3f} ms, in-wave exchange {inwave_ms:.3f} ms "
          f"(same bits: {inwave_same})", file=sys.stderr, flush=True)
    ref.copy_(y)
    e_plain = max(errors(xs, ref[: sample * 64].cpu().numpy().reshape(sample, 16, 4), sample))
    rows = {}
    for s in (6, 7, 8, 12, 15, 16):
        ms = timed(kbsrc_amd.split_dft16, s)
        e = max(errors(xs, y[: sample * 64].cpu().numpy().reshape(sample, 16, 4), sample))
        d = float(((y - ref).abs().max() / ref.abs().max()).item())
        # int8 matrix work: per 16 vectors, S (S + 1) / 2 digit pairs x 2 M-tiles of 16 x 16 x 32 MACs (2 ops each)
        ops = (nvec / 16) * (s * (s + 1) / 2) * 2 * 16 * 16 * 32 * 2
        rows[s] = {"ms": round(ms, 4), "ps_per_point": round(ms * 1e9 / (16 * nvec), 2),
                   "err_vs_256bit_log2": round(float(np.log2(e)), 1), "diff_vs_plain_dd": d,
                   "speed_vs_plain": round(plain_ms / ms, 3),
                   "int8_tops": round(ops / (ms / 1e3) / 1e12, 1),
                   "int8_frac_of_dense_peak": round(ops / (ms / 1e3) / INT8_PEAK_OPS, 3)}
        print(f"slices {s}: {ms:.3f} ms, error 2^{np.log2(e):.1f}, x{plain_ms / ms:.3f} vs plain", file=sys.stderr,
              flush=True)
    ok = [s for s in rows if rows[s]["err_vs_256bit_log2"] <= -100]
    s100 = min(ok) if ok else None
    del x, y, ref
    torch.cuda.empty_cache()
    conv = bench_split_conversion(args, (6, 7, 8, s100 or 16))
    line = {
        "metric": "split-MFMA vs plain dd radix-16 DFT speed ratio at <= 2^-100 error",
        "value": rows[s100]["speed_vs_plain"] if s100 else None,
        "unit": "x (plain dd time / split time)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": rows[s100]["ms"] if s100 else None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int8 MFMA + f64",
        "data": "synthetic",
        "config": {"workload": f"{nvec} forward DFTs of 16 complex dd values (67.1 M points), magnitudes 2^-20..2^20",
                   "slices_at_2^-100": s100, "parallelism": "single GPU"},
        "plain_dd": {"ms": round(plain_ms, 4), "ps_per_point": round(plain_ms * 1e9 / (16 * nvec), 2),
                     "err_vs_256bit_log2": round(float(np.log2(e_plain)), 1)},
        "plain_dd_inwave_exchange": {"ms": round(inwave_ms, 4), "speed_vs_lds_exchange": round(plain_ms / inwave_ms, 3),
                                     "same_bits": inwave_same},
        "split": rows,
        "conversion": conv,
    }
    print(json.dumps(line), flush=True)
    return 0


def _block_reference(xx, info, rate_in, rate_out, channel=0):
    """A middle block of one channel and its 2^-256 restatement (tests/hp_block.py): (block index, first output
    sample of the block's kept range, fixed-point reference samples, the block's peak as a double)."""
    import numpy as np

    import kbsrc_amd
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import hp_block as hb
    up = bool(info["up"])
    b = info["nblk"] // 2
    P, Q, hop = info["P"], info["Q"], info["hop"]
    base = channel * info["xstride"]
    blk = xx[base + b * hop * Q: base + b * hop * Q + info["Nb"]].cpu().numpy()
    fk = rate_out if up else rate_in
    gc = math.gcd(info["c2"], fk)
    taps = kbsrc_amd.kernel_taps(info["Lk"], info["c2"] // gc, fk // gc, 40.0)
    spec = kbsrc_amd.kernel_spectrum(taps, info["Mb"] if up else info["Nb"], info["Nb"])
    ref = hb.block_output(blk, spec, P, Q, info["nb"], info["marg"], up)
    return b, b * hop * P - info["o"], ref, float(np.abs(hb.to_double(ref)).max())


def _pair_error(pairs, channel, nout, d0, ref, peak):
    """log2 of the largest |hi + lo - ref| over the peak, for output samples d0.. of one channel of a pair buffer."""
    import hp_block as hb
    p = pairs[2 * (channel * nout + d0): 2 * (channel * nout + d0 + len(ref))].cpu().numpy().reshape(-1, 2)
    diff = hb.fix(p[:, 0]) + hb.fix(p[:, 1]) - ref
    err = max(math.ldexp(float(abs(int(v))), -hb.F) for v in diff) / peak
    return round(math.log2(err), 1) if err > 0 else None


def internal_errors(plan, xx, info, stream, slices, rate_in, rate_out=RATE_OUT):
    """The internal dd error, before the final rounding, of the plain form and the split form at `slices`, over one
    middle block of channel 0 against the 2^-256 block restatement (tests/hp_block.py), each over the block's peak
    (log2): what "at 2^-100" refers to. kbsrc_plan_execute_dd writes the samples as (hi, lo) pairs."""
    import torch
    _, d0, ref, peak = _block_reference(xx, info, rate_in, rate_out)
    pairs = torch.empty(2 * CH * info["nout"], dtype=torch.float64, device=xx.device)
    out = {}
    for name, s in (("plain", 0), (f"split_{slices}", slices)):
        plan.execute_dd(xx.data_ptr(), pairs.data_ptr(), s, stream)
        torch.cuda.synchronize()
        out[name] = _pair_error(pairs, 0, info["nout"], d0, ref, peak)
    del pairs
    return out


def bench_split_conversion(args, slices):
    """Config 3's conversions, stereo 180 s exact at the automatic block: 96 kHz -> 44.1 kHz (its 320-point forward
    row sub-FFTs take the radix-16 matrix stage) and 44.1 kHz -> 48 kHz (its 224-point inverse column sub-FFTs), plain
    and in the split form at each slice count: time per conversion, the largest difference from the plain output
    over every sample, and on one middle block against the 2^-256 restatement (tests/hp_block.py) both the error of
    the f64 output and the internal dd error before the final rounding (kbsrc_plan_execute_dd)."""
    import numpy as np
    import torch

    import kbsrc_amd
    import hp_block as hb
    from tests.helpers import input_44k_180s, input_96k_180s

    def one(x, rate_in, rate_out):
        frames = x.shape[0]
        plan = kbsrc_amd.Plan(frames, CH, rate_in, rate_out, "exact")
        info = plan.info
        dev = torch.device("cuda:0")
        xx = torch.empty(CH * info["xstride"], dtype=torch.float64, device=dev)
        out = torch.empty(CH * info["nout"], dtype=torch.float64, device=dev)
        plain = torch.empty_like(out)
        pairs = torch.empty(2 * CH * info["nout"], dtype=torch.float64, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        plan.stage(x, xx.data_ptr(), stream)

        def timed(run):
            for _ in range(max(1, args.warmup)):
                run()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                run()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) * 1e3 / args.steps

        plain_ms = timed(lambda: plan.execute(xx.data_ptr(), plain.data_ptr(), stream))
        peak_all = float(plain.abs().max().item())
        t0 = time.perf_counter()
        b, d0, ref, peak = _block_reference(xx, info, rate_in, rate_out)

        def f64_err(y):
            return round(math.log2(float(hb.errors(y[d0:d0 + len(ref)].cpu().numpy(), ref).max()) / peak), 1)

        def dd_err(s):
            plan.execute_dd(xx.data_ptr(), pairs.data_ptr(), s, stream)
            torch.cuda.synchronize()
            return _pair_error(pairs, 0, info["nout"], d0, ref, peak)

        e_plain, i_plain = f64_err(plain), dd_err(0)
        name = f"stereo {rate_in} -> {rate_out} Hz, 180 s, exact, block {info['nb']}"
        print(f"{name}: plain {plain_ms:.3f} ms; block {b}: f64 output error 2^{e_plain} of peak, internal dd error "
              f"2^{i_plain} (2^-256 block in {time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
        rows = {}
        for sl in slices:
            ms = timed(lambda: plan.execute_split(xx.data_ptr(), out.data_ptr(), sl, stream))
            d = float((out - plain).abs().max().item()) / peak_all
            rows[sl] = {"ms": round(ms, 3), "speed_vs_plain": round(plain_ms / ms, 3),
                        "max_diff_vs_plain_over_peak_log2": round(math.log2(d), 1) if d > 0 else None,
                        "block_err_vs_2^-256_over_peak_log2": f64_err(out),
                        "internal_dd_err_vs_2^-256_over_peak_log2": dd_err(sl)}
            print(f"  split S={sl}: {ms:.3f} ms (x{plain_ms / ms:.3f}), vs plain 2^{rows[sl]['max_diff_vs_plain_over_peak_log2']}, "
                  f"f64 error 2^{rows[sl]['block_err_vs_2^-256_over_peak_log2']}, internal dd error "
                  f"2^{rows[sl]['internal_dd_err_vs_2^-256_over_peak_log2']}", file=sys.stderr, flush=True)
        plan.close()
        del xx, out, plain, pairs
        torch.cuda.empty_cache()
        return {"workload": name,
                "plain": {"ms": round(plain_ms, 3), "block_err_vs_2^-256_over_peak_log2": e_plain,
                          "internal_dd_err_vs_2^-256_over_peak_log2": i_plain},
                "split": rows}

    res = [one(input_96k_180s(RATE_IN * SECONDS), RATE_IN, RATE_OUT)]
    res.append(one(input_44k_180s(), 44100, 48000))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", default="exact", choices=["exact", "fast"])
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="reference CPU engine threads beside its default (0: min(16, usable cores))")
    ap.add_argument("--cpu-runs", type=int, default=5, help="alternating CPU/GPU runs of the baseline protocol")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-split", action="store_true", help="skip timing the split-precision form")
    ap.add_argument("--profile-copy", action="store_true",
                    help="also run a plain 256 MiB device copy (calibrates HBM byte counters under rocprofv3 --pmc)")
    ap.add_argument("--batch-set", default="long", choices=["long", "short"],
                    help="batch workload: long = 64 jobs of 0.5-300 s, 1/2/6/8 channels (default); short = 256 jobs "
                         "of 0.5-2 s, 1/2 channels")
    ap.add_argument("--tune", action="store_true",
                    help="measure kernel shapes once before timing (Plan.tune / Batch.tune; same output bits)")
    ap.add_argument("--workload", default="single", choices=["single", "batch", "split", "extreme"],
                    help="single: the stereo 96 kHz -> 44.1 kHz 180 s job (default); batch: 64 mixed jobs in one "
                         "batch; split: int8-MFMA split-precision radix-16 DFTs against the plain dd pass; extreme: "
                         "10 MHz -> 48 kHz 30 s and 8-channel 384 kHz -> 44.1 kHz 600 s at --len 0.5")
    ap.add_argument("--summarize-split", metavar="DIR",
                    help="print the radix-16 stage summary of a rocprofv3 kernel trace of the split workload in DIR")
    ap.add_argument("--summarize-counters", metavar="DIR",
                    help="print the per-pass VALU/LDS counter summary of rocprofv3 csv output under DIR/{pmc1,pmc2}")
    ap.add_argument("--summarize-kernels", metavar="DIR",
                    help="print the per-kernel summary (calls, durations) of rocprofv3 --stats csv output in DIR")
    ap.add_argument("--summarize-rocprof", metavar="DIR",
                    help="print the per-pass summary of rocprofv3 csv output under DIR/{stats,fetch,write} and exit")
    args = ap.parse_args()
    if args.summarize_rocprof:
        return summarize_rocprof(args.summarize_rocprof, args.mode, args.steps, args.warmup)
    if args.summarize_counters:
        return summarize_counters(args.summarize_counters, args.mode)
    if args.summarize_split:
        return summarize_split(args.summarize_split)
    if args.summarize_kernels:
        return summarize_kernels(args.summarize_kernels)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): the path would shard by block ranges with "
              "no exchange step, and nothing here drives a second GPU; run with --gpus 1", file=sys.stderr)
        return 2

    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "split":
        return bench_split(args)
    if args.workload == "extreme":
        return bench_extreme(args)

    import numpy as np
    import torch

    import kbsrc_amd
    from tests.helpers import input_96k_180s

    frames = RATE_IN * SECONDS
    x = input_96k_180s(frames)
    t0 = time.perf_counter()
    plan = kbsrc_amd.Plan(frames, CH, RATE_IN, RATE_OUT, args.mode)
    t_design = time.perf_counter() - t0
    info = plan.info
    dev = torch.device("cuda:0")
    xx = torch.empty(CH * info["xstride"], dtype=torch.float64, device=dev)
    out = torch.empty(CH * info["nout"], dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    t0 = time.perf_counter()
    plan.stage(x, xx.data_ptr(), stream)
    torch.cuda.synchronize()
    t_stage = time.perf_counter() - t0
    print(f"plan: nb={info['nb']} Nb={info['Nb']} Mb={info['Mb']} tasks={info['tasks']} "
          f"design {t_design:.3f}s stage {t_stage:.3f}s", file=sys.stderr, flush=True)

    if args.tune:
        t0 = time.perf_counter()
        plan.tune(xx.data_ptr(), out.data_ptr(), stream)
        print(f"tuned {kbsrc_amd.tuned_shapes()} pass shapes in {time.perf_counter() - t0:.2f} s", file=sys.stderr,
              flush=True)
    if args.profile_copy:
        a = torch.ones(1 << 25, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    # untimed warm-up: W executes, and at least WARM_MIN in all, so that the GPU's clocks have left their idle state
    # (after a pause of the GPU the first ~5 executes of this workload run 10-25% slower; rocprofv3 trace)
    warm = max(args.warmup, WARM_MIN)
    for _ in range(warm):
        plan.execute(xx.data_ptr(), out.data_ptr(), stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.execute(xx.data_ptr(), out.data_ptr(), stream)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    nout_total = CH * info["nout"]
    value = nout_total / (ms / 1e3) / 1e6
    print(f"engine {ms:.3f} ms/step -> {value:.1f} output Msamples/s ({warm} warm-up executes)", file=sys.stderr,
          flush=True)

    # per-kernel HIP-event times on the launch stream right after the timed executes (the five launches queued back
    # to back, timing-only events read after), averaged over a few executes
    reps = 5
    acc = None
    for _ in range(reps):
        t = plan.execute_timed(xx.data_ptr(), out.data_ptr(), stream)
        acc = t if acc is None else {k: acc[k] + t[k] for k in acc}
    kt = {k: v / reps for k, v in acc.items()}
    if kt["mid"] == 0 and kt["fwd_row"] > 0:   # the forward rows and the split/fold ran as one kernel
        kt = {"fwd_col": kt["fwd_col"], "row_mid": kt["fwd_row"], "inv_col": kt["inv_col"], "inv_row": kt["inv_row"]}

    # what was timed is checked: the output of the timed executes (the timed-pass executes rewrite the same bits)
    # against the reference's pinned SHA-256 of the f64 WAV it writes for this input
    from tests.helpers import PINS_96K_180S, wav_sha_planar
    host = out.cpu().numpy().reshape(CH, info["nout"])
    sha_ok = wav_sha_planar(host, RATE_OUT) == PINS_96K_180S[f"{args.mode}_f64"]
    del host
    print(f"timed output sha256 equals the reference pin: {sha_ok}", file=sys.stderr, flush=True)
    tasks = info["tasks"]
    fl = block_flops(info["Nb"], info["Mb"], args.mode == "exact", bool(info["up"]))
    passes = {}
    for k in kt:
        tf = pass_flops(fl, k) * tasks / (kt[k] / 1e3) / 1e12 if kt[k] > 0 else 0.0
        passes[k] = {"ms": round(kt[k], 4), "gflop": round(pass_flops(fl, k) * tasks / 1e9, 3), "tflops": round(tf, 2)}
    dom = max(kt, key=kt.get)
    dd = args.mode == "exact"
    by = block_bytes(info["Nb"], info["Mb"], dd, info["hop"] * info["P"])
    for k in kt:
        nbytes = by[k] * tasks + (k in ("mid", "row_mid")) * (info["Nb"] // 2 + 1) * (16 if dd else 8)
        passes[k]["algorithmic_gb"] = round(nbytes / 1e9, 4)
        passes[k]["gbs"] = round(nbytes / (kt[k] / 1e3) / 1e9, 1) if kt[k] > 0 else 0.0
    prof_us = committed_avg_us(args.mode, dom)
    if dd:
        # dd arithmetic: ~12 flops per compulsory byte in the dominant pass, above the FP64 ridge (78.6 TFLOP/s over
        # 8 TB/s = 9.8 flop/B): bound by FP64 instruction issue on the vector ALUs
        achieved = pass_flops(fl, dom) * tasks / (kt[dom] / 1e3) / 1e12
        roof = {"bound": "fp64-valu", "kernel": dom, "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4)}
        if prof_us:
            roof["frac_from_profile"] = round(pass_flops(fl, dom) * tasks / (prof_us / 1e6) / 1e12 / FP64_PEAK_TFLOPS, 4)
    else:
        # FP64 arithmetic: ~1.4 flops per compulsory byte, far below the ridge: HBM bound
        achieved = passes[dom]["gbs"]
        roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4)}
    roof["avg_launch_us_events"] = round(kt[dom] * 1e3, 1)
    roof["avg_launch_us_profile"] = prof_us
    roof["algorithmic_per_launch"] = (f"{pass_flops(fl, dom) * tasks / 1e9:.3f} GFLOP" if dd
                                      else f"{passes[dom]['algorithmic_gb']:.4f} GB")
    roof["traffic"] = committed_traffic(args.mode, dom)
    # the same pass's VALU issue rate (SQ_ACTIVE_INST_VALU against one instruction per 4 cycles per SIMD): dd adds
    # count one flop each against a peak that counts an FMA as two, so frac stays well under this
    roof["valu_busy"] = committed_valu_busy(args.mode, dom)

    # split-precision form of the same conversion at the first slice count whose radix-16 stage error is within
    # 2^-100 (15 slices: 2^-103.8 against a 256-bit DFT), timed like the plain form, outside its timed region
    split = None
    if dd and not args.no_split:
        alt = torch.empty_like(out)
        for _ in range(warm):
            plan.execute_split(xx.data_ptr(), alt.data_ptr(), SPLIT_SLICES_2M100, stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.execute_split(xx.data_ptr(), alt.data_ptr(), SPLIT_SLICES_2M100, stream)
        torch.cuda.synchronize()
        split_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        diff = float((alt - out).abs().max().item())
        split = {"slices": SPLIT_SLICES_2M100, "ms_per_step": round(split_ms, 4),
                 "ratio": round(ms / split_ms, 3), "max_abs_diff_vs_plain": diff}
        print(f"split form, {SPLIT_SLICES_2M100} slices: {split_ms:.3f} ms/step (x{ms / split_ms:.3f} of plain), "
              f"max |diff| vs plain {diff:.3g}", file=sys.stderr, flush=True)
        del alt
        t0 = time.perf_counter()
        split["internal_error_over_peak_log2"] = internal_errors(plan, xx, info, stream, SPLIT_SLICES_2M100, RATE_IN)
        print(f"internal dd error before the final rounding, one block against 2^-256 (log2 of peak): "
              f"{split['internal_error_over_peak_log2']} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr,
              flush=True)

    graph_ms, graph_same = graph_timing(args, plan, xx, out)
    print(f"captured graph replay: {graph_ms:.3f} ms/step; same output: {graph_same}", file=sys.stderr, flush=True)

    # the clock the timed region ran at: power and GFX clocks while the same execute repeats back to back
    power = power_sample(lambda: plan.execute(xx.data_ptr(), out.data_ptr(), stream), torch.cuda.synchronize)
    if power:
        print(f"under load: socket power {power['socket_power_w']:.0f} W, GFX clock {power['gfx_clock_mhz']:.0f} MHz "
              f"(XCDs {min(power['gfx_clock_mhz_per_xcd']):.0f}-{max(power['gfx_clock_mhz_per_xcd']):.0f})",
              file=sys.stderr, flush=True)
        if dd:
            # the FP64 peak scales with the clock: at the clock the power limit leaves, the attainable peak is lower
            at = FP64_PEAK_TFLOPS * power["gfx_clock_mhz"] / PEAK_CLOCK_MHZ
            roof["peak_at_measured_clock"] = round(at, 2)
            roof["frac_at_measured_clock"] = round(roof["achieved"] / at, 4)
            # the step's VALU issue floor at that clock: every VALU wave-instruction of an execute (committed
            # counters) at one per 4 cycles on each of the 1024 SIMDs
            insts = committed_valu_insts(args.mode)
            if insts:
                floor_ms = insts * 4 / 1024 / (power["gfx_clock_mhz"] * 1e6) * 1e3
                roof["step_valu_issue_floor_ms_at_measured_clock"] = round(floor_ms, 4)
                roof["step_frac_of_issue_floor"] = round(floor_ms / ms, 4)
        roof["under_load"] = {k: power[k] for k in ("socket_power_w", "gfx_clock_mhz")}

    # CPU baseline protocol: the reference's own engine (oracle/_ref/kbsrc, built from its sources) on the same input
    # on this host, at its default thread count for the mode (8 exact, 4 fast: src/resampler.cpp:139) and at the
    # CPU share one GPU gets on this host (16 threads; the reference's useful cap here is ch * ceil(nblk / 4) = 54,
    # src/resampler.cpp:208), alternating with the GPU's three measured forms: (a) device kernels (the engine
    # over HBM-resident input), (b) the engine with the input's upload and staging and the output's download to
    # pinned host memory, (c) kbsrc_resample host memory to host memory (plan, design, upload, staging, engine,
    # download). Median of the runs each.
    cpu = None
    ref = os.path.join(ROOT, "oracle", "_ref", "kbsrc")
    runs = args.cpu_runs
    gpu_runs = {"a_kernels_ms": [], "b_engine_h2d_d2h_ms": [], "c_host_to_host_ms": []}
    settings = {}
    have_ref = not args.no_cpu and os.path.exists(ref)
    default_threads = 8 if dd else 4
    share = args.cpu_threads or min(16, host_cpu()["usable_cpus"])
    thread_list = sorted({default_threads, share})
    pinned = torch.empty(nout_total, dtype=torch.float64, pin_memory=True)
    from tests.helpers import wav_f64
    with tempfile.TemporaryDirectory(prefix="kbsrc-bench-") as td:
        src, dst = os.path.join(td, "in.wav"), os.path.join(td, "out.wav")
        if have_ref:
            with open(src, "wb") as f:
                f.write(wav_f64(x, CH, RATE_IN))
            ref_engine_once(src, dst, share, extra=("--mode", args.mode))   # warm-up (page cache, first run)
        for i in range(runs):
            if have_ref:
                for th in thread_list:
                    r = ref_engine_once(src, dst, th, extra=("--mode", args.mode))
                    if r is None:
                        have_ref = False
                        break
                    settings.setdefault(th, []).append(r)
                    print(f"cpu reference run {i + 1}/{runs}: {th} threads, engine {r[0]:.3f} s, design {r[1]:.3f} s",
                          file=sys.stderr, flush=True)
            for _ in range(WARM_MIN):   # the GPU idled during the CPU runs: clocks up again first
                plan.execute(xx.data_ptr(), out.data_ptr(), stream)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(5):
                plan.execute(xx.data_ptr(), out.data_ptr(), stream)
            torch.cuda.synchronize()
            gpu_runs["a_kernels_ms"].append((time.perf_counter() - t0) * 1e3 / 5)
            t0 = time.perf_counter()
            plan.stage(x, xx.data_ptr(), stream)
            plan.execute(xx.data_ptr(), out.data_ptr(), stream)
            pinned.copy_(out, non_blocking=True)
            torch.cuda.synchronize()
            gpu_runs["b_engine_h2d_d2h_ms"].append((time.perf_counter() - t0) * 1e3)
            t0 = time.perf_counter()
            kbsrc_amd.resample(x, RATE_IN, RATE_OUT, args.mode)
            gpu_runs["c_host_to_host_ms"].append((time.perf_counter() - t0) * 1e3)
            print(f"gpu run {i + 1}/{runs}: (a) {gpu_runs['a_kernels_ms'][-1]:.3f} ms, (b) "
                  f"{gpu_runs['b_engine_h2d_d2h_ms'][-1]:.2f} ms, (c) {gpu_runs['c_host_to_host_ms'][-1]:.2f} ms",
                  file=sys.stderr, flush=True)
    del pinned
    gpu_med = {k: round(median(v), 3) for k, v in gpu_runs.items()}
    if have_ref and settings:
        hc = host_cpu()
        per = {}
        for th, rs in sorted(settings.items()):
            eng = median([r[0] for r in rs])
            per[str(th)] = {"engine_s": round(eng, 4), "design_s": round(median([r[1] for r in rs]), 4),
                            "engine_output_msamples_s": round(nout_total / eng / 1e6, 2),
                            "speedup_a_kernels": round(eng * 1e3 / gpu_med["a_kernels_ms"], 1),
                            "speedup_b_engine_h2d_d2h": round(eng * 1e3 / gpu_med["b_engine_h2d_d2h_ms"], 1),
                            "speedup_c_host_to_host": round(eng * 1e3 / gpu_med["c_host_to_host_ms"], 1),
                            "runs": len(rs)}
        best = min(settings, key=lambda th: median([r[0] for r in settings[th]]))
        eng = median([r[0] for r in settings[best]])
        cpu = {"value": round(nout_total / eng / 1e6, 2), "unit": "output Msamples/s", "cores": best,
               "kind": "reference",
               "sample": f"full workload (2 x {frames} frames), the reference's engine (resample time of its -v line, "
                         f"design and file I/O excluded), median of {runs} runs alternating with the GPU runs, at "
                         f"{' and '.join(str(t) for t in thread_list)} threads; value at {best}",
               "cpu_model": hc["model"], "host_logical_cpus": hc["logical_cpus"], "usable_cpus": hc["usable_cpus"],
               "threads": per}

    algo_gflop = sum(fl.values()) * tasks / 1e9
    line = {
        "metric": "exact-mode output Msamples/s (96k->44.1k)" if args.mode == "exact"
        else "fast-mode output Msamples/s (96k->44.1k)",
        "value": round(value, 2),
        "unit": "output Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_executes_run": warm,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "stereo 96000->44100 Hz, 180 s float64 uniform noise, beta 40, len 0.2 s, LPC on, "
                               f"{args.mode} mode", "frames_per_channel": frames, "channels": CH,
                   "nb": info["nb"], "tasks": tasks, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "output_sha_ok": bool(sha_ok),
        "passes": passes,
        "pass_sum_ms": round(sum(kt.values()), 4),
        "algorithmic_gflop_per_step": round(algo_gflop, 2),
        "step_tflops": round(algo_gflop / ms, 3),
        "design_s": round(t_design, 3),
        "stage_s": round(t_stage, 3),
        "gpu_forms_median_ms": gpu_med,
        "graph_ms_per_step": round(graph_ms, 4),
        "graph_outputs_identical": graph_same,
        "tuned": bool(args.tune),
    }
    if split:
        line["split_vs_plain_at_2^-100"] = split["ratio"]
        line["split"] = split
    if cpu:
        line["speedup_vs_cpu_engine"] = round(value / cpu["value"], 1)
    plan.close()
    report_children()
    print(json.dumps(line), flush=True)
    return 0


def power_sample(run, sync, seconds=3.0):
    """Socket power and GFX clocks while `run` (one execute) repeats back to back: one `amd-smi metric` reading
    taken 1.5 s into the loop. The published FP64 peak assumes the 2.4 GHz peak clock; under the board's power limit
    the clock settles lower, and the attainable peak with it. None when amd-smi is absent or unreadable."""
    import threading
    smi = "/opt/rocm/bin/amd-smi"
    if not os.path.exists(smi):
        return None
    got = {}

    def sample():
        time.sleep(1.5)
        try:
            r = subprocess.run([smi, "metric", "-c", "-p"], capture_output=True, text=True, timeout=30)
            got["text"] = r.stdout
        except (OSError, subprocess.SubprocessError) as e:
            got["error"] = str(e)

    th = threading.Thread(target=sample)
    th.start()
    t_end = time.time() + seconds
    n = 0
    while th.is_alive() or time.time() < t_end:
        for _ in range(20):
            run()
        sync()
        n += 20
    th.join()
    r = parse_amd_smi_metric(got.get("text", ""))
    if r:
        r["executes_in_loop"] = n
    return r


def parse_amd_smi_metric(text):
    """Socket power (W) and the GFX clocks (MHz, one per XCD) from `amd-smi metric -c -p` text, or None."""
    pw = re.search(r"SOCKET_POWER:\s*([0-9.]+)\s*W", text)
    # the CLOCK section lists GFX_0 .. GFX_7 (one per XCD), each with its current CLK first
    clks = [float(m.group(1)) for m in re.finditer(r"GFX_\d+:\s*\n\s*CLK:\s*([0-9.]+)\s*MHz", text)]
    if not pw or not clks:
        return None
    return {"socket_power_w": float(pw.group(1)), "gfx_clock_mhz": round(sum(clks) / len(clks), 1),
            "gfx_clock_mhz_per_xcd": clks}


def report_children():
    """Names any child process of this one still alive at the end (a bench must leave none)."""
    kids = []
    try:
        for t in os.listdir(f"/proc/{os.getpid()}/task"):
            with open(f"/proc/{os.getpid()}/task/{t}/children") as f:
                kids += f.read().split()
    except OSError:
        return
    for k in kids:
        try:
            with open(f"/proc/{k}/cmdline", "rb") as f:
                cmd = f.read().replace(b"\0", b" ").decode(errors="replace")
        except OSError:
            cmd = "?"
        print(f"child process still alive at exit: {k} {cmd}", file=sys.stderr, flush=True)


if __name__ == "__main__":
    sys.exit(main())
