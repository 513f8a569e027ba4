"""The bench's readers of tool output, on recorded text (no GPU): amd-smi's power and clock report."""
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
