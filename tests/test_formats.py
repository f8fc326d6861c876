"""Host-side formats and generators: the reference's snapshot/.init layout,
its input-parameter file, the grids and the seeded initial field."""
import math

import numpy as np

from channel_rl_amd import cases, snapshot


def test_snapshot_round_trip(tmp_path):
    c, (U, V, W) = cases.small_case(10, 9, 12, seed=1)
    P = np.random.default_rng(0).standard_normal((c.nxg, c.nyg, c.nzg))
    path = str(tmp_path / "s.init")
    snapshot.write(path, c.x, c.y, c.z, U, V, W, P, P * 2)
    r = snapshot.read(path, with_extra=True)
    for name, a in (("U", U), ("V", V), ("W", W), ("P", P), ("nu_t", 2 * P)):
        assert np.array_equal(r[name], a)
    assert np.array_equal(r["xm"], 0.5 * (c.x[1:] + c.x[:-1]))


def test_snapshot_layout_matches_the_reference_reader(tmp_path):
    # scripts/check_file.m: int32 n + float64 mesh arrays, then int32[3] dims per field
    c, (U, V, W) = cases.small_case(10, 9, 12, seed=1)
    path = str(tmp_path / "s.init")
    snapshot.write(path, c.x, c.y, c.z, U, V, W)
    raw = open(path, "rb").read()
    n = np.frombuffer(raw[:4], "<i4")[0]
    assert n == 10
    off = 4 + 8 * 10
    assert np.frombuffer(raw[off:off + 4], "<i4")[0] == 9
    mesh = 4 * 6 + 8 * (10 + 9 + 12 + 9 + 8 + 11)
    assert tuple(np.frombuffer(raw[mesh:mesh + 12], "<i4")) == (10, 10, 13)
    total = mesh + 3 * 12 + 8 * (U.size + V.size + W.size)
    assert len(raw) == total


def test_parameter_file_format(tmp_path):
    p = tmp_path / "input_parameters"
    snapshot.write_parameters(str(p), 34, 33, 34, 0.5, 1 / 180, 1.0, 1, 2, 10, 11, 5, 5, "a.init", "out")
    lines = p.read_text().splitlines()
    assert len(lines) == 22 and lines[0].startswith("!") and lines[1].split() == ["34", "33", "34"]


def test_grids():
    y = cases.tanh_y(33, 1.5)
    assert y[0] == 0.0 and y[-1] == 2.0 and np.all(np.diff(y) > 0)
    assert np.allclose(y, 2.0 - y[::-1], atol=1e-14)
    j = np.arange(33)
    assert np.allclose(y, 1 + np.tanh(1.5 * (2 * j / 32 - 1)) / math.tanh(1.5), atol=1e-15)
    x = cases.periodic_faces(34, 2 * math.pi)
    assert math.isclose(x[-2], 2 * math.pi) and math.isclose(x[-1] - x[-2], x[1] - x[0])


def test_noise_is_seeded_and_bounded():
    a = cases.uniform_noise(1234, 1, (5, 6, 7))
    b = cases.uniform_noise(1234, 1, (5, 6, 7))
    c = cases.uniform_noise(1235, 1, (5, 6, 7))
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    big = cases.uniform_noise(7, 2, (64, 64, 64))
    assert big.min() >= -1.0 and big.max() < 1.0 and abs(big.mean()) < 0.01


def test_bulk_velocity_correlation():
    assert 15.0 < cases.bulk_velocity_plus(180.0) < 16.5


def test_rocprof_summary_names_and_traffic_table(tmp_path):
    from channel_rl_amd import rocprof_summary as rs
    assert rs.short_name("void chan::k_rhs_rk<double, 64, 4>(double const*, chan::Dims)") == "k_rhs_rk<double, 64, 4>"
    assert rs.short_name("(anonymous namespace)::k_stream_copy(HIP_vector_type<double, 2u> const*, long)") == "k_stream_copy"
    assert rs.base_name("chan::k_fftz<-1>(chan::Dims, chan::PoissonBufs)") == "k_fftz"
    hdr = "Correlation_Id,Kernel_Name,Counter_Name,Counter_Value\n"
    f = tmp_path / "f.csv"
    w = tmp_path / "w.csv"
    f.write_text(hdr + '1,"(anonymous namespace)::k_stream_copy(x)",FETCH_SIZE,524288\n'
                 '2,"void chan::k_dsm_fused<double, 30, 6>(x)",FETCH_SIZE,100\n')
    w.write_text(hdr + '1,"(anonymous namespace)::k_stream_copy(x)",WRITE_SIZE,1048576\n'
                 '2,"void chan::k_dsm_fused<double, 30, 6>(x)",WRITE_SIZE,10\n')
    res, sf, sw = rs.traffic(str(f), str(w))
    assert sf == 2.0 and sw == 1.0
    assert res["k_dsm_fused<double, 30, 6>"][3] == 100 * 1024 * 2 + 10 * 1024
    t = tmp_path / "t.md"
    t.write_text(rs.traffic_table(str(f), str(w)))
    assert rs.load_traffic(str(t))["k_dsm_fused"] == 100 * 1024 * 2 + 10 * 1024


def test_committed_traffic_table_parses():
    import os
    from channel_rl_amd import rocprof_summary as rs
    p = os.path.join(os.path.dirname(__file__), "..", "profiles", "hbm_traffic_channel_step_192x97x192_fp64.md")
    t = rs.load_traffic(p)
    assert t["k_dsm_fused"] > 0 and t["k_rhs_rk"] > 0


def test_rocprof_summary_sq_fractions(tmp_path):
    # one kernel of 2 waves over 10 cycles per XCD: every derived column by hand
    from channel_rl_amd import rocprof_summary as rs
    hdr = "Kernel_Name,Grid_Size,Counter_Name,Counter_Value\n"
    vals = {"SQ_WAVE_CYCLES": 100, "SQ_WAIT_ANY": 40, "SQ_WAIT_INST_ANY": 20, "SQ_ACTIVE_INST_ANY": 40,
            "SQ_ACTIVE_INST_VALU": 1280, "SQ_LDS_IDX_ACTIVE": 1280, "SQ_LDS_BANK_CONFLICT": 64,
            "SQ_INSTS_VALU": 600, "SQ_INSTS_LDS": 30, "GRBM_GUI_ACTIVE": 80}
    p = tmp_path / "sq.csv"
    p.write_text(hdr + "".join(f'"void chan::k_dsm_fused<double, 30, 14>(x)",128,{c},{v}\n' for c, v in vals.items()))
    rows = [r for r in rs.sq_table([str(p)]).splitlines() if r.startswith("| k_")]
    assert rows == ["| k_dsm_fused<double, 30, 14> | 2 | 300 | 15 | 0.40 | 0.20 | 0.40 | 0.50 | 0.50 | 0.05 |"]


def test_step_algorithmic_bytes_of_the_dynamic_model_step():
    """bench.py sums each stage's algorithmic values per launch over the
    launches of a step: the dynamic-model step with mass-flow mode 1 (pass A,
    RHS and the projection's five kernels per substep, pass B in the last
    substep only, compute_dt and the shift folded) moves 107 values per cell,
    856 B in FP64 and 548 B with the FP32 bulk (spectral values stay FP64)."""
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location(
        "bench_mod", os.path.join(os.path.dirname(__file__), "..", "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    per_step = {"k_dsm_fused": 3, "k_rhs_rk": 3, "k_dsm_nut": 1, "k_fftx_fwd": 3, "k_fftz": 6,
                "k_tridiag": 3, "k_fftx_inv": 3, "k_project": 3}

    def nbytes(bulk):
        return sum(n * (bench.ALG_VALUES[k][0] * bulk + bench.ALG_VALUES[k][1] * 8) for k, n in per_step.items())

    assert sum(n * sum(bench.ALG_VALUES[k]) for k, n in per_step.items()) == 107
    assert nbytes(8) == 856
    assert nbytes(4) == 548
