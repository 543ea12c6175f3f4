"""Host-side constants of bench.py that its roofline line is computed from.

The algorithmic bytes per column are the reference data flow's compulsory
I/O at 127 layers (LW 128 g-points, SW 112 g-points): the gas stage's
FP32 outputs follow from the array shapes, and the stage figures add up to
the per-scheme totals of that data flow (LW 551,816 B, SW 399,796 B per
column: gas, cloud, solver and heating-rate stages).
"""
import bench

NLAY, G_LW, G_SW = 127, 128, 112


def test_alg_bytes_are_whole_bytes():
    for k, v in bench.ALG_BYTES_PER_COLUMN.items():
        assert isinstance(v, int) and v > 10_000, k


def test_gas_outputs_follow_from_shapes():
    b = bench.ALG_BYTES_PER_COLUMN
    # LW: tau, lay_src, lev_src_inc, lev_src_dec (layer x g) + sfc_src, sfc_src_Jac (g)
    lw_out = 4 * NLAY * G_LW * 4 + 2 * G_LW * 4
    # SW: tau, ssa, g (layer x g) + toa_src (g)
    sw_out = 3 * NLAY * G_SW * 4 + G_SW * 4
    assert b["lw_gas_optics"] - lw_out == 4_584
    assert b["sw_gas_optics"] - sw_out == 4_576


def test_stage_bytes_add_to_scheme_totals():
    b = bench.ALG_BYTES_PER_COLUMN
    assert b["lw_gas_optics"] + b["lw_cloud_optics"] + b["lw_lw_solve"] == 551_816
    assert b["sw_gas_optics"] + b["sw_cloud_optics"] + b["sw_sw_solve"] == 399_796
    # the solvers' heating-rate share: p_lev and fluxes in, FP64 rate out
    assert b["lw_lw_solve"] - (269_308 + 2_048) == 3_072 + 1_016 == 8 * (3 * (NLAY + 1)) + 8 * NLAY


def test_traffic_table_covers_the_roofline_kernels():
    for k in ("sw_sw_solve", "lw_lw_solve", "lw_gas_optics", "sw_gas_optics"):
        assert 1.0 < bench.TRAFFIC_GB[k] < 100.0, k
    assert len(bench.TRAFFIC_SRC_SHA) == 16
