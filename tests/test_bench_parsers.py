"""The bench's readers of tool output, on recorded text (no GPU): amd-smi's power and clock report, the pass labels of
profiled dispatches and the fused pass's flop and byte model."""
import bench

# `amd-smi metric -c -p` on an MI355X under load (GFX_2 .. GFX_6 elided), as the bench reads it
SMI = """GPU: 0
    POWER:
        SOCKET_POWER: 1381 W
        GFX_VOLTAGE: N/A
        SOC_VOLTAGE: N/A
        MEM_VOLTAGE: N/A
        THROTTLE_STATUS: N/A
        POWER_MANAGEMENT: N/A
    CLOCK:
        GFX_0:
            CLK: 2123 MHz
            MIN_CLK: 500 MHz
            MAX_CLK: 2400 MHz
            CLK_LOCKED: DISABLED
            DEEP_SLEEP: ENABLED
        GFX_1:
            CLK: 2061 MHz
            MIN_CLK: 500 MHz
            MAX_CLK: 2400 MHz
            CLK_LOCKED: DISABLED
            DEEP_SLEEP: ENABLED
        GFX_7:
            CLK: 1979 MHz
            MIN_CLK: 500 MHz
            MAX_CLK: 2400 MHz
            CLK_LOCKED: DISABLED
            DEEP_SLEEP: ENABLED
        MEM_0:
            CLK: 2000 MHz
            MIN_CLK: 2000 MHz
            MAX_CLK: 2000 MHz
"""


def test_amd_smi_power_and_gfx_clocks():
    r = bench.parse_amd_smi_metric(SMI)
    assert r["socket_power_w"] == 1381.0
    assert r["gfx_clock_mhz_per_xcd"] == [2123.0, 2061.0, 1979.0]   # the memory clock is not a GFX clock
    assert r["gfx_clock_mhz"] == round((2123 + 2061 + 1979) / 3, 1)


def test_amd_smi_unusable_text():
    assert bench.parse_amd_smi_metric("") is None
    assert bench.parse_amd_smi_metric("GPU: 0\n    POWER:\n        SOCKET_POWER: N/A\n") is None


def test_dispatch_labels_of_both_engine_forms():
    """Profiled dispatches are labelled by pass: five per execute with the split/fold kernel, four where the forward
    row pass and the split/fold run as one kernel; a split-form execute (another row kernel) stays unlabelled."""
    col = "void kbsrc::gpu::(anonymous namespace)::k_fixed<kbsrc::gpu::DDS, R<3,3,5,7>, 4, 256, 0, 0, true>(PassArgs)"
    row = "void kbsrc::gpu::(anonymous namespace)::k_fixed<kbsrc::gpu::DDS, R<4,4,4,5>, 4, 256, 1, 0, true>(PassArgs)"
    mid = "void kbsrc::gpu::(anonymous namespace)::k_mid<kbsrc::gpu::DDA>(MidArgs)"
    rowmid = "void kbsrc::gpu::(anonymous namespace)::k_rowmid<kbsrc::gpu::DDS, R<4,4,4,5>, 315, 315, 2, 256, false>(...)"
    inv_c = "void kbsrc::gpu::(anonymous namespace)::k_fixed<kbsrc::gpu::DDS, R<3,3,3,7>, 8, 256, 1, 0, true>(PassArgs)"
    inv_r = "void kbsrc::gpu::(anonymous namespace)::k_fixed<kbsrc::gpu::DDS, R<5,7,7>, 5, 256, 1, 2, true>(PassArgs)"
    split = "void kbsrc::gpu::(anonymous namespace)::k_fixed_split<kbsrc::gpu::DDS, R<4,4,4,5>, 4, 256, 15>(PassArgs)"
    names = [col, row, mid, inv_c, inv_r, col, rowmid, inv_c, inv_r, "copyBuffer", col, split, mid, inv_c, inv_r]
    assert bench._pass_of_dispatches(names) == [
        "fwd_col", "fwd_row", "mid", "inv_col", "inv_row", "fwd_col", "row_mid", "inv_col", "inv_row", None,
        None, None, None, None, None]


def test_fused_pass_model():
    """The fused pass's flops are the row pass's and the split/fold's; its bytes are the column output read and the
    inverse input written (the forward spectrum never leaves LDS)."""
    fl = bench.block_flops(201600, 92610, True, False)
    by = bench.block_bytes(201600, 92610, True, 74382)
    assert bench.pass_flops(fl, "row_mid") == fl["fwd_row"] + fl["mid"]
    assert sorted(fl) == ["fwd_col", "fwd_row", "inv_col", "inv_row", "mid"]   # the step's total sums these once
    assert by["row_mid"] == 32 * (100800 + 46305)
    assert by["row_mid"] < by["fwd_row"] + by["mid"]
