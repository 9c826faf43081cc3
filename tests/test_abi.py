"""C ABI and host-side checks that need no GPU: the library builds and loads,
exports every entry point include/dsgdpp.h declares, refuses to run without a
device (no silent CPU fallback), the CLIs parse like the reference tool, and
the Python mirror renders the reference's functor strings."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dsgdpp.h")
LIB = os.path.join(ROOT, "dsgdpp_amd", "libdsgdpp_amd.so")
CLI = os.path.join(ROOT, "dsgdpp_amd", "bin", "mfdsgdpp")
ORACLE_CLI = os.path.join(ROOT, "oracle", "mfdsgdpp_cpu")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(dsgdpp_[a-z_]+)\s*\(", txt)) - {"dsgdpp_cfg", "dsgdpp_run_cfg"})


def test_header_declarations_exported():
    lib = ctypes.CDLL(LIB)
    names = declared()
    assert "dsgdpp_epoch" in names and "dsgdpp_create" in names
    for n in names:
        assert hasattr(lib, n), n
    from dsgdpp_amd import _lib
    assert sorted(_lib.EXPORTS) == names


def _gpu_count():
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        c = ctypes.c_int(0)
        return c.value if hip.hipGetDeviceCount(ctypes.byref(c)) == 0 else 0
    except OSError:
        return 0


@pytest.mark.skipif(_gpu_count() > 0, reason="checks the no-device error path")
def test_create_without_device_fails_loudly():
    import dsgdpp_amd as D
    with pytest.raises(D.DsgdPpError) as e:
        D.DsgdPp(rank=4, tasks_per_rank=2)
    assert e.value.code == -4


def test_invalid_configurations_rejected_before_device():
    import dsgdpp_amd as D
    for kw, code in [(dict(update="Nzsl_L2(1)", loss="Nzsl"), -1),           # mfdsgdpp.cc:353
                     (dict(update="Foo", loss="Foo"), -1),
                     (dict(regularize="L2(0.1)"), -1),
                     (dict(update="Nzsl_L2()", loss="Nzsl_L2()"), -1),
                     (dict(truncate="(1)"), -1),
                     (dict(rank=0), -1),
                     (dict(rank=300), -5),
                     (dict(stratum_order=7), -1)]:
        with pytest.raises(D.DsgdPpError) as e:
            D.DsgdPp(**{**dict(rank=4, tasks_per_rank=2), **kw})
        assert e.value.code == code, kw


def test_cli_requires_options():
    r = subprocess.run([CLI], capture_output=True, text=True)
    assert r.returncode == 1 and "have to be present" in r.stderr
    r = subprocess.run([CLI, "--input-file", "a", "--input-row-file", "b", "--input-col-file", "c", "--update",
                        "Nzsl", "--regularize", "None", "--loss", "Nzsl", "--rank", "2", "--decay", "Const(1)",
                        "--abs"], capture_output=True, text=True)
    assert r.returncode == 1 and "Invalid combination" in r.stderr
    r = subprocess.run([CLI, "--input-file", "a", "--input-row-file", "b", "--input-col-file", "c", "--update",
                        "Nzsl", "--regularize", "None", "--loss", "Nzsl", "--rank", "2", "--decay", "Const(1)",
                        "--sgd-order", "XYZ"], capture_output=True, text=True)
    assert r.returncode == 1 and "sgdOrder" in r.stderr


def _write_mm(tmp, m, n, r, seed=0):
    rng = np.random.default_rng(seed)
    cells = rng.choice(m * n, size=40, replace=False)
    with open(tmp / "v.mmc", "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n% test\n")
        f.write(f"{m} {n} {len(cells)}\n")
        for c in cells:
            f.write(f"{c // n + 1} {c % n + 1} {rng.normal():.17g}\n")
    for name, a, b in (("w.mma", m, r), ("h.mma", r, n)):
        with open(tmp / name, "w") as f:
            f.write(f"%%MatrixMarket matrix array real general\n{a} {b}\n")
            for v in rng.uniform(-0.5, 0.5, a * b):
                f.write(f"{v:.17g}\n")


def test_oracle_cli_trace_and_outputs(tmp_path):
    _write_mm(tmp_path, 10, 12, 3)
    r = subprocess.run([ORACLE_CLI, "--input-file", str(tmp_path / "v.mmc"), "--input-row-file",
                        str(tmp_path / "w.mma"), "--input-col-file", str(tmp_path / "h.mma"),
                        "--output-row-file", str(tmp_path / "wo.mma"), "--output-col-file", str(tmp_path / "ho.mma"),
                        "--trace", str(tmp_path / "t.R"), "--update", "Nzsl_L2(0.1)", "--regularize", "None",
                        "--loss", "Nzsl_L2(0.1)", "--rank", "3", "--decay", "BoldDriver(0.05)", "--epochs", "4",
                        "--tasks-per-rank", "2", "--seed", "3"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    t = (tmp_path / "t.R").read_text()
    assert t.startswith("trace <- list(trace=list(\nlist(it=0, epoch=0, loss=") and t.endswith("));")
    assert t.count("list(it=") == 5 and 'update = "Nzsl_L2(0.1)"' in t
    lines = (tmp_path / "wo.mma").read_text().splitlines()
    assert lines[0] == "%%MatrixMarket matrix array real general" and lines[3] == "10 3" and len(lines) == 34


def test_mmc_errors_reported(tmp_path):
    from oracle import pyoracle as po
    (tmp_path / "bad.mmc").write_text("%%MatrixMarket matrix array real general\n2 2\n")
    o = po.Oracle(rank=2, tasks_per_rank=1)
    import ctypes as C
    rc = po.lib().orc_load_mmc(o._h, 0, str(tmp_path / "bad.mmc").encode())
    assert rc != 0 and b"banner" in po.lib().orc_last_error(o._h)


def test_python_mirror_specs():
    import dsgdpp_amd as D
    assert D.UpdateNzslL2(1).spec() == "Nzsl_L2(1)"
    assert D.UpdateNzslL2(0.05).spec() == "Nzsl_L2(0.05)"
    assert D.SumLoss(D.NzslLoss(), D.L2Loss(1)).spec() == "Nzsl_L2(1)"
    assert D.SumLoss(D.NzslLoss(), D.Nzl2Loss(2)).spec() == "Nzsl_Nzl2(2)"
    assert D.BoldDriver(0.001).spec() == "BoldDriver(0.001)"
    assert D.BoldDriver(0.5, 0.25, 1.1).spec() == "BoldDriver(0.5,0.25,1.1)"
    b = D.BoldDriver(1.0)
    assert b(None, 5.0) == 1.0 and b(5.0, 4.0) == 1.05 and b(4.0, 4.0) == 0.525
    assert D.UpdateTruncate(D.UpdateNzsl(), -1, 2).truncate() == "(-1,2)"


def test_mfdsgd_cli_options():
    """bin/mfdsgd is the DSGD tool (tools/mfdsgd.cc:50-233): its usage names
    it and --map-reduce; mfdsgdpp rejects --map-reduce; the oracle CLI's
    --algorithm DSGD runs the d x d restatement."""
    dsgd = os.path.join(ROOT, "dsgdpp_amd", "bin", "mfdsgd")
    r = subprocess.run([dsgd], capture_output=True, text=True)
    assert r.returncode == 1 and "mfdsgd [options]" in r.stderr and "--map-reduce" in r.stderr
    base = ["--input-file", "a", "--input-row-file", "b", "--input-col-file", "c", "--update", "Nzsl",
            "--regularize", "None", "--loss", "Nzsl", "--rank", "2", "--decay", "Const(1)"]
    r = subprocess.run([dsgd] + base + ["--map-reduce", "maybe"], capture_output=True, text=True)
    assert r.returncode == 1 and "map-reduce" in r.stderr
    r = subprocess.run([CLI] + base + ["--map-reduce", "true"], capture_output=True, text=True)
    assert r.returncode == 1 and "unrecognised option '--map-reduce'" in r.stderr


def test_oracle_cli_dsgd(tmp_path):
    _write_mm(tmp_path, 10, 12, 3)
    args = ["--input-file", str(tmp_path / "v.mmc"), "--input-row-file", str(tmp_path / "w.mma"),
            "--input-col-file", str(tmp_path / "h.mma"), "--update", "Nzsl_L2(0.1)", "--regularize", "None",
            "--loss", "Nzsl_L2(0.1)", "--rank", "3", "--decay", "BoldDriver(0.05)", "--epochs", "2",
            "--tasks-per-rank", "3", "--seed", "3"]
    r = subprocess.run([ORACLE_CLI] + args + ["--algorithm", "DSGD", "--output-row-file", str(tmp_path / "a.mma")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r2 = subprocess.run([ORACLE_CLI] + args + ["--output-row-file", str(tmp_path / "b.mma")],
                        capture_output=True, text=True)
    assert r2.returncode == 0, r2.stderr
    # d x d and d x 2d blocking visit the points in other orders
    assert (tmp_path / "a.mma").read_text() != (tmp_path / "b.mma").read_text()
    r = subprocess.run([ORACLE_CLI] + args + ["--algorithm", "PSGD"], capture_output=True, text=True)
    assert r.returncode == 1 and "algorithm" in r.stderr
