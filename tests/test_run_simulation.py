"""run_simulation.py drop-in: CLI, CSV format and naming, the engine-process
protocol (run_simulation.py:59-256), and -- on a GPU -- the in-process and
shim engines against the oracle."""
import json
import os
import stat
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN_DIR, PKG_DIR, channel_params

MATRIX = os.path.join(GOLDEN_DIR, "AR4JA_r45_4c_128c_r12.chinn.out")
RUN = os.path.join(PKG_DIR, "run_simulation.py")
sys.path.insert(0, PKG_DIR)
import run_simulation as RS  # noqa: E402


def _args(*extra):
    return RS.build_parser().parse_args(["--chinn", MATRIX, "--ebn0_start", "1.0", "--ebn0_end", "2.0", *extra])


def test_output_filename_rule():
    # expected strings follow run_simulation.py:112-137 of the reference
    assert RS.output_filename(_args()) == os.path.join(
        "Results", "results_AR4JA_r45_4c_128c_r12_amin_ebn0_1.00_2.00_step0.25_max100000_min50_iter16_440cores.csv")
    a = _args("--nms", "--single_core", "--l_max", "8", "--r_max", "6", "--v_max", "20", "--ebn0_step", "0.1",
              "--max_blocks_per_core", "7", "--min_errors", "25", "--max_iter", "200")
    assert RS.output_filename(a) == os.path.join(
        "Results", "results_AR4JA_r45_4c_128c_r12_nms_clipL8.0R6.0V20.0_ebn0_1.00_2.00_step0.10_max7_min25"
                   "_iter200_1core.csv")
    assert RS.output_filename(_args("--engine", "circulant")).endswith("_amin_circulant_ebn0_1.00_2.00_step0.25"
                                                                       "_max100000_min50_iter16_440cores.csv")
    assert RS.output_filename(_args("--output", "x.csv")) == os.path.join("Results", "x.csv")


def test_csv_rows():
    assert RS.CSV_HEADER == "EbN0_dB,Total_Blocks,Total_Bit_Errors,Total_Block_Errors,BER,FER\n"
    assert RS.csv_row(1.5, 1000, 37, 2, 2048) == "1.50,1000,37,2,1.806641e-05,2.000000e-03\n"
    assert RS.csv_row(2.0, 0, 0, 0, 2048) == "2.00,0,0,0,0.000000e+00,0.000000e+00\n"
    assert RS.csv_row_circulant(2.5, 4096, 10, 1, 28000, 2.0) == \
        "2.50,4096,10,1,2.384186e-06,2.441406e-04,6.836,4.19,2.00\n"


def test_parse_matrix_header():
    assert RS.parse_matrix_header(MATRIX) == (2560, 1536, 6, 6)
    with pytest.raises(FileNotFoundError):
        RS.parse_matrix_header("/nonexistent.chinn.out")


def _fake_engine(tmp_path, rc=0, final=True):
    """An engine binary that records its argv and speaks the JSON-lines protocol."""
    p = tmp_path / "fake_engine"
    p.write_text(f"""#!{sys.executable}
import json, sys
open({str(tmp_path / 'argv.log')!r}, 'a').write(' '.join(sys.argv[1:]) + '\\n')
e = float(sys.argv[sys.argv.index('-e') + 1])
print('[CONFIG] fake engine')
print(json.dumps({{"progress": True, "ebn0": e, "blocks": 440, "max_blocks": 880, "bit_errors": 3,
                  "frame_errors": 1, "target_errors": 5, "ber": 0.0, "fer": 0.0, "throughput_msps": 1.0,
                  "elapsed_sec": 0.1}}))
if {final!r}:
    print(json.dumps({{"ebn0": e, "blocks": 880, "bit_errors": int(e * 10), "frame_errors": 2, "elapsed_sec": 0.5}}))
sys.exit({rc})
""")
    p.chmod(p.stat().st_mode | stat.S_IEXEC)
    return str(p)


def test_binary_protocol_and_csv(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    eng = _fake_engine(tmp_path)
    out = RS.main(["--chinn", MATRIX, "--ebn0_start", "1.0", "--ebn0_end", "1.5", "--ebn0_step", "0.5",
                   "--punctured", "512", "--binary", eng, "--nms", "--l_max", "4", "--max_iter", "50",
                   "--min_errors", "5", "--max_blocks", "880"])
    lines = open(out).read().splitlines()
    assert lines[0] == RS.CSV_HEADER.strip()
    assert lines[1] == "1.00,880,10,2,5.548651e-06,2.272727e-03"
    assert lines[2] == "1.50,880,15,2,8.322976e-06,2.272727e-03"
    argv = open(tmp_path / "argv.log").read().splitlines()
    assert argv[0].split() == ["-c", MATRIX, "-e", "1.0", "-p", "512", "-t", "5", "-k", "2", "-i", "50", "-n",
                               "-b", "1000", "-L", "4.0"]


def test_binary_failure_stops_the_sweep(tmp_path, monkeypatch, capsys):
    monkeypatch.chdir(tmp_path)
    out = RS.main(["--chinn", MATRIX, "--ebn0_start", "1.0", "--ebn0_end", "2.0", "--ebn0_step", "0.5",
                   "--binary", _fake_engine(tmp_path, rc=3)])
    assert open(out).read() == RS.CSV_HEADER
    assert "exited with code 3" in capsys.readouterr().err
    out = RS.main(["--chinn", MATRIX, "--ebn0_start", "1.0", "--ebn0_end", "1.0",
                   "--binary", _fake_engine(tmp_path, final=False), "--output", "nofinal.csv"])
    assert open(out).read() == RS.CSV_HEADER


def test_combine_script_columns(tmp_path, monkeypatch):
    """The columns combine_and_plot_results.py reads (:99-115): SER and FER
    are computable from our CSV exactly as it computes them."""
    import pandas as pd
    monkeypatch.chdir(tmp_path)
    out = RS.main(["--chinn", MATRIX, "--ebn0_start", "1.0", "--ebn0_end", "1.0",
                   "--binary", _fake_engine(tmp_path)])
    df = pd.read_csv(out)
    ser = df["Total_Bit_Errors"] / (df["Total_Blocks"] * 2560.0)
    assert list(df.columns[:6]) == ["EbN0_dB", "Total_Blocks", "Total_Bit_Errors", "Total_Block_Errors", "BER",
                                    "FER"]
    assert ser.iloc[0] == pytest.approx(10 / (880 * 2560.0)) and df["FER"].iloc[0] == pytest.approx(2 / 880)


# ----------------------------------------------------------------- GPU ---

@pytest.mark.gpu
def test_inprocess_sweep_matches_oracle(tmp_path, monkeypatch, oracle):
    from ldpc_mi355x import codes
    monkeypatch.chdir(tmp_path)
    print("[gpu-case] run_simulation in-process", flush=True)
    out = RS.main(["--chinn", MATRIX, "--ebn0_start", "1.2", "--ebn0_end", "1.6", "--ebn0_step", "0.4",
                   "--punctured", "512", "--max_blocks", "2200", "--batch_size", "2", "--max_iter", "30",
                   "--min_errors", "100000"])
    M, N, rows = codes.read_chinn_out(MATRIX)
    lines = open(out).read().splitlines()[1:]
    for line, ebn0 in zip(lines, (1.2, 1.6)):
        mu, sg = channel_params(ebn0)
        o = oracle.run_point_r(N, rows, 2048, 133742, 0, 2200, mu, sg, max_iter=30)
        assert line == RS.csv_row(ebn0, o["blocks"], o["bit_errors"], o["frame_errors"], 2048).strip()


@pytest.mark.gpu
def test_shim_binary_matches_inprocess(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    shim = os.path.join(PKG_DIR, "ldpc_mi355x", "ldpc_sim_scalar")
    common = ["--chinn", MATRIX, "--ebn0_start", "1.0", "--ebn0_end", "1.5", "--ebn0_step", "0.5", "--punctured",
              "512", "--max_blocks", "3080", "--batch_size", "3", "--max_iter", "25", "--min_errors", "8", "--nms"]
    print("[gpu-case] run_simulation --binary ldpc_sim_scalar vs in-process", flush=True)
    a = open(RS.main(common + ["--output", "a.csv"])).read()
    b = open(RS.main(common + ["--output", "b.csv", "--binary", shim])).read()
    assert a == b


@pytest.mark.gpu
def test_circulant_engine_csv(tmp_path, monkeypatch, oracle, code12):
    monkeypatch.chdir(tmp_path)
    print("[gpu-case] run_simulation --engine circulant", flush=True)
    out = RS.main(["--chinn", MATRIX, "--ebn0_start", "1.5", "--ebn0_end", "1.5", "--punctured", "512",
                   "--engine", "circulant", "--max_blocks", "4400", "--batch_size", "5", "--max_iter", "200",
                   "--min_errors", "100000"])
    row = open(out).read().splitlines()[1].split(",")
    mu, sg = channel_params(1.5)
    o = oracle.run_point_t(code12, 133742, 0, 4400, mu, sg, 200, 0.9375)
    assert row[:4] == ["1.50", "4400", str(o["bit_errors"]), str(o["frame_errors"])]
    assert row[6] == f"{o['iter_sum'] / 4400:.3f}"


@pytest.mark.gpu
def test_circulant_shim_writes_engine_csv(tmp_path, oracle, code12):
    shim = os.path.join(PKG_DIR, "ldpc_mi355x", "ldpc_sim_circulant")
    print("[gpu-case] ldpc_sim_circulant", flush=True)
    p = subprocess.run([shim, "1.5", "3520", "200", "--single-core", "--chunk-batches", "440", "--min-fe", "1000",
                        "--seed", "9", "--out", str(tmp_path / "c.csv")], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    rows = open(tmp_path / "c.csv").read().splitlines()
    assert rows[0].startswith("EbN0_dB,Total_Blocks,Total_Bit_Errors,Total_Block_Errors,BER,FER,Mean_Iterations")
    assert len(rows) == 3 and rows[-1].split(",")[1] == "3520"   # two chunks of 1760, cumulative rows
    mu, sg = channel_params(1.5)
    o = oracle.run_point_t(code12, 9, 0, 3520, mu, sg, 200, 0.9375)
    assert rows[-1].split(",")[2:4] == [str(o["bit_errors"]), str(o["frame_errors"])]


@pytest.mark.gpu
def test_circulant_engine_random_info_csv(tmp_path, monkeypatch, oracle, code12):
    monkeypatch.chdir(tmp_path)
    print("[gpu-case] run_simulation --engine circulant --random_info", flush=True)
    out = RS.main(["--chinn", MATRIX, "--ebn0_start", "1.5", "--ebn0_end", "1.5", "--punctured", "512",
                   "--engine", "circulant", "--random_info", "--max_blocks", "4400", "--batch_size", "5",
                   "--max_iter", "200", "--min_errors", "100000", "--output", "r.csv"])
    row = open(out).read().splitlines()[1].split(",")
    mu, sg = channel_params(1.5)
    o = oracle.run_point_t_cw(code12, 133742, 0, 4400, mu, sg, 200, 0.9375)
    assert row[:4] == ["1.50", "4400", str(o["bit_errors"]), str(o["frame_errors"])]


def test_random_info_needs_the_inprocess_circulant_engine(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    with pytest.raises(SystemExit):
        RS.main(["--chinn", MATRIX, "--ebn0_start", "1.5", "--ebn0_end", "1.5", "--random_info", "--binary",
                 "/bin/false"])
