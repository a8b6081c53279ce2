"""The C-ABI library (CPU only: loads it, no GPU calls)."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

import timbir_amd as tb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in ("mbir4d.h", "timbir_dev.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b([a-z_][a-z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name.startswith("tb_") or name == "reconstruct":
                names.append(name)
    return sorted(set(names))


def test_library_exports_every_declared_entry_point():
    lib = tb.lib()
    names = declared_functions()
    assert "reconstruct" in names and "tb_stage_create" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(set(tb.EXPORTS)) == names


def header_struct(name):
    """(type, field, count) of `typedef struct { ... } name;` in timbir_dev.h;
    count is None for a scalar, else the array length."""
    text = open(os.path.join(ROOT, "include", "timbir_dev.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    consts = {k: int(v) for k, v in re.findall(r"\b(TB_\w+)\s*=\s*(\d+)", text)}
    body = re.search(r"typedef struct \{([^}]*)\}\s*" + name + ";", text).group(1)
    out = []
    for m in re.finditer(r"(\w+)\s+(\w+)(?:\[(\w+)\])?;", body):
        t, f, n = m.groups()
        out.append((t, f, None if n is None else int(consts.get(n, n))))
    return out


@pytest.mark.parametrize("name,cls", [("tb_options", "TbOptions"), ("tb_stats", "TbStats")])
def test_binding_structs_match_the_header(name, cls):
    # the ctypes mirrors have the header's fields in the header's order and
    # types, so the library and the binding agree on every offset
    ctype = {"int32_t": C.c_int32, "uint32_t": C.c_uint32, "int64_t": C.c_int64,
             "double": C.c_double}
    fields = header_struct(name)
    got = getattr(tb, cls)._fields_
    assert [f for _, f, _ in fields] == [f for f, _ in got]
    want = [ctype[t] if n is None else ctype[t] * n for t, _, n in fields]
    assert [C.sizeof(t) for t in want] == [C.sizeof(t) for _, t in got]
    assert [getattr(t, "_type_", t) for t in want] == [getattr(t, "_type_", t) for _, t in got]


def test_options_round_trip_without_gpu():
    o = tb.get_options()
    assert o.max_iterations == 1000
    tb.set_options(order_seed=7)
    assert tb.get_options().order_seed == 7
    tb.set_options(order_seed=0)


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(tb, "_LIB", None)
    monkeypatch.setattr(tb, "LIB_PATH", "/nonexistent/libtimbir_hip.so")
    with pytest.raises(tb.TimbirError):
        tb.lib()


DRIVER = os.path.join(ROOT, "timbir_amd", "tb_bin_data")


def test_bin_data_driver_rejects_missing_input(tmp_path):
    """The bin_data driver reads its inputs before any GPU call: a missing
    projections.bin ends it with status 1 and the reference's message."""
    assert os.path.exists(DRIVER), "build() makes timbir_amd/tb_bin_data"
    r = subprocess.run([DRIVER, "--proj_rows", "4", "--proj_cols", "16", "--proj_num", "8",
                        "--recon_num", "1", "--vox_wid", "1", "--rot_center", "8"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "ERROR: read_BinFile" in r.stdout
    assert "huber_delta = 0.500000, huber_T = 4.000000" in r.stdout  # bin_data/XT_Main.h


def test_bin_data_driver_rejects_bad_arguments(tmp_path):
    r = subprocess.run([DRIVER, "--proj_rows", "4", "extra"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1
    assert "ERROR: read_command_line_args" in r.stdout


def test_bench_refuses_more_than_one_gpu():
    # the bench exits with a message, before any device work, for N > 1
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stdout


def test_bench_traffic_comes_from_the_committed_counter_summary():
    # the roofline's HBM traffic per launch is read from the counter summary
    # in profiles/ (FETCH_SIZE + WRITE_SIZE, copy-calibrated); it must parse
    # and exceed the kernel's algorithmic bytes per launch (1.85e10 B)
    sys.path.insert(0, ROOT)
    import bench
    t = bench.measured_traffic()
    assert t is not None and t > 1.85e10
    assert os.path.exists(bench.TRAFFIC_PROFILE)
