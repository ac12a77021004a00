"""Host code under AddressSanitizer and UBSan (CPU only).

* the oracle (oracle/piro_oracle.c) driven through every restated function
  on small and degenerate grids by oracle/sanitize_check.c;
* the library's host input/output (piro_amd/csrc/host_io.cpp: the .ini
  reader and the multi-threaded Paraview writer) driven by
  piro_amd/csrc/sanitize/host_io_check.cpp: it must write the bytes the
  reference's own writer wrote for the frozen fields (tests/golden/ref_vtk)
  and read the reference's case files as the driver restatement does.
Any sanitizer report fails the run (halt on the first error, leaks checked).
"""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import case as pc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
HOST_IO = os.path.join(ROOT, "piro_amd", "csrc", "sanitize", "_build", "host_io_check")
ORACLE_CHECK = os.path.join(ROOT, "oracle", "_build", "sanitize_check")


@pytest.fixture(scope="module")
def built():
    for d, target in ((os.path.join(ROOT, "oracle"), "sanitize"), (os.path.join(ROOT, "piro_amd", "csrc", "sanitize"), "all")):
        r = subprocess.run(["make", "-s", "-C", d, target], capture_output=True, text=True)
        if r.returncode != 0:
            pytest.fail(f"sanitizer build failed in {d}:\n{r.stderr[-2000:]}")


def run(args, **kw):
    r = subprocess.run(args, capture_output=True, text=True, env=ENV, **kw)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]
    return r


def test_oracle_under_asan_ubsan(built):
    r = run([ORACLE_CHECK])
    assert r.returncode == 0, r.stderr[-3000:]
    assert r.stdout.strip().endswith("sanitize_check ok")


def vtk_cases():
    with open(os.path.join(GOLD, "ref_digests.json")) as f:
        return json.load(f)["vtk"]


def write_with_host_io(tmp_path, dims, fields, n_scal, n_vec, step):
    inp = tmp_path / "in.bin"
    np.concatenate(fields).astype("<f4").tofile(inp)
    out = tmp_path / "out"
    r = run([HOST_IO, "vtk", *map(str, dims), str(n_scal), str(n_vec), str(step), str(inp), str(out)])
    assert r.returncode == 0, r.stderr[-3000:]
    tag = "%05d" % step
    base = out / "case" / "mesh"
    return (base / "level" / f"{tag}_level_0.vti").read_bytes(), (base / f"mesh_{tag}.vth").read_bytes()


@pytest.mark.parametrize("meta", vtk_cases(), ids=lambda m: m["key"])
def test_paraview_writer_under_asan_matches_reference_bytes(built, tmp_path, meta):
    d = os.path.join(GOLD, "ref_vtk", meta["key"])
    z = np.load(os.path.join(d, "fields.npz"))
    fields = [z[f"f{k}"] for k in range(meta["n_scal"] + meta["n_vec"])]
    vti, vth = write_with_host_io(tmp_path, meta["dims"], fields, meta["n_scal"], meta["n_vec"], meta["step"])
    tag = "%05d" % meta["step"]
    with open(os.path.join(d, f"{tag}_level_0.vti"), "rb") as f:
        assert vti == f.read()
    with open(os.path.join(d, f"mesh_{tag}.vth"), "rb") as f:
        assert vth == f.read()


def test_paraview_writer_under_asan_many_slabs(built, tmp_path):
    # 38 interior z-slabs: every writer thread formats a slab range
    dims = (7, 6, 40)
    N = 7 * 6 * 40
    rng = np.random.default_rng(11)
    fields = [(rng.standard_normal(N) * 10.0 ** rng.integers(-6, 7, N)).astype(np.float32),
              rng.standard_normal(3 * N).astype(np.float32)]
    vti, vth = write_with_host_io(tmp_path, dims, fields, 1, 1, 321)
    typed = [("S0", 1, fields[0]), ("V0", 2, fields[1])]
    assert vti == (pc.writevti(dims, (0, 0, 0), (1, 1, 1), typed) + "\n").encode()
    assert vth == (pc.writevth(dims, 321) + "\n").encode()


@pytest.mark.parametrize("name", ["HeatEq.ini", "AdvEq.ini", "LUDecomp.ini", "setup.ini", "solver.ini",
                                  "IC/distribution.ini", "Data/hashtable.ini"])
def test_ini_reader_under_asan(built, name):
    path = os.path.join(GOLD, "cases", name)
    table = pc.read_ini(path)
    pairs = [(sec, key) for sec, kv in table.items() for key in kv] + [("Mesh", "no_such_key"), ("Nope", "n")]
    args = [HOST_IO, "ini", path]
    for sec, key in pairs:
        args += [sec, key]
    r = run(args)
    assert r.returncode == 0, r.stderr[-3000:]
    got = r.stdout.split("\n")[:-1]
    want = [table.get(sec, {}).get(key, "default_value") for sec, key in pairs]
    assert got == want
