"""Pins of the case-driver restatement (oracle/case.py) and of the host-only
Paraview entry point, on the CPU.

* the oracle's Paraview text equals, byte for byte, what the reference's own
  writer (paraview::writevth/writevti + bc::prepbc, compiled from the PIRO
  sources) wrote for the fixtures in tests/golden/ref_vtk;
* piro_paraview_write (libpiro_hip, host code) writes the same bytes;
* the reference's case files parse to the values their text states.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle as po
from oracle import case as pc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
CASES = os.path.join(GOLD, "cases")


def vtk_cases():
    with open(os.path.join(GOLD, "ref_digests.json")) as f:
        return json.load(f)["vtk"]


def load_case(meta):
    d = os.path.join(GOLD, "ref_vtk", meta["key"])
    z = np.load(os.path.join(d, "fields.npz"))
    fields = [z[f"f{k}"] for k in range(meta["n_scal"] + meta["n_vec"])]
    tag = "%05d" % meta["step"]
    with open(os.path.join(d, f"{tag}_level_0.vti"), "rb") as f:
        vti = f.read()
    with open(os.path.join(d, f"mesh_{tag}.vth"), "rb") as f:
        vth = f.read()
    return fields, vti, vth


@pytest.mark.parametrize("meta", vtk_cases(), ids=lambda m: m["key"])
def test_oracle_writer_matches_reference_bytes(meta):
    fields, vti, vth = load_case(meta)
    names = meta["names"][1:]
    typed = [(names[k], 1 if k < meta["n_scal"] else 2, fields[k]) for k in range(len(fields))]
    n = tuple(meta["dims"])
    # writefile appends a newline (src/fileutilities.cpp:24-40)
    assert (pc.writevti(n, (0, 0, 0), (1, 1, 1), typed) + "\n").encode() == vti
    assert (pc.writevth(n, meta["step"]) + "\n").encode() == vth


@pytest.mark.parametrize("meta", vtk_cases(), ids=lambda m: m["key"])
def test_library_writer_matches_reference_bytes(meta, tmp_path):
    from piro_amd._lib import check, lib
    fields, vti, vth = load_case(meta)
    names = meta["names"]                      # constant "Alpha" first, as the case driver passes it
    types = [0] + [1 if k < meta["n_scal"] else 2 for k in range(len(fields))]
    arrays = [np.full(int(np.prod(meta["dims"])), 0.1, np.float32)] + [np.ascontiguousarray(f) for f in fields]
    c_names = (C.c_char_p * len(names))(*[s.encode() for s in names])
    c_types = (C.c_int * len(types))(*types)
    c_ptrs = (C.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])
    dims = (C.c_int * 3)(*meta["dims"])
    check(lib.piro_paraview_write(str(tmp_path).encode(), b"case", meta["step"], dims, None, None, len(arrays),
                                  c_names, c_types, c_ptrs))
    tag = "%05d" % meta["step"]
    base = tmp_path / "case" / "mesh"
    assert (base / "level" / f"{tag}_level_0.vti").read_bytes() == vti
    assert (base / f"mesh_{tag}.vth").read_bytes() == vth


def test_library_writer_rejects_bad_arguments(tmp_path):
    from piro_amd._lib import lib
    dims = (C.c_int * 3)(2, 4, 4)
    assert lib.piro_paraview_write(str(tmp_path).encode(), b"c", 0, dims, None, None, 0, None, None, None) == 1
    dims = (C.c_int * 3)(4, 4, 4)
    assert lib.piro_paraview_write(str(tmp_path).encode(), b"c", -1, dims, None, None, 0, None, None, None) == 1


@pytest.mark.parametrize("dims", [(5, 6, 3), (21, 18, 67)])
def test_library_writer_values_on_many_slabs(dims, tmp_path):
    # a random field with special values, over enough z-slabs that the
    # writer splits it between threads: the array text is the interior
    # cells, x fastest, each spelled as the reference's stream spells a
    # float, joined by single spaces in cell order
    from piro_amd._lib import check, lib
    nx, ny, nz = dims
    rng = np.random.default_rng(nz)
    f = (rng.standard_normal(nx * ny * nz) * 10.0 ** rng.integers(-30, 30, nx * ny * nz)).astype(np.float32)
    inner = f.reshape(nz, ny, nx)[1:-1, 1:-1, 1:-1]  # a view: the specials go to interior cells
    for k, v in enumerate([0.0, -0.0, np.inf, -np.inf, np.nan, -np.nan, 1e-45, 3.4e38, 1e5]):
        inner[np.unravel_index(k, inner.shape)] = v
    names = (C.c_char_p * 1)(b"T")
    types = (C.c_int * 1)(1)
    ptrs = (C.c_void_p * 1)(f.ctypes.data)
    check(lib.piro_paraview_write(str(tmp_path).encode(), b"c", 7, (C.c_int * 3)(*dims), None, None, 1, names,
                                  types, ptrs))
    text = (tmp_path / "c" / "mesh" / "level" / "00007_level_0.vti").read_text()
    body = text.split('format="ascii">\n', 1)[1].split("</DataArray>", 1)[0]
    g = f.reshape(nz, ny, nx)[1:-1, 1:-1, 1:-1].ravel()
    assert body == " ".join(pc.fmt_float(v) for v in g)
    assert "-inf" in body and "-nan" in body and "1.4013e-45" in body


def test_fmt_float_spellings():
    cases = [(0.0, "0"), (-0.0, "-0"), (1.0 / 3, "0.333333"), (123456789.0, "1.23457e+08"), (1e-5, "1e-05"),
             (100000.0, "100000"), (1e6, "1e+06"), (float("inf"), "inf"), (float("-inf"), "-inf")]
    for v, s in cases:
        assert pc.fmt_float(np.float32(v)) == s
    assert pc.fmt_float(np.float32("nan")) == "nan"
    assert pc.fmt_float(-np.float32("nan")) == "-nan"


def test_ini_reader_rules(tmp_path):
    p = tmp_path / "a.ini"
    p.write_text("; comment\n# also\n[S]\n  key  =  a  b  \nx=1\n[T]\nx = 2\nbad line\n")
    d = pc.read_ini(str(p))
    assert d == {"S": {"key": "a  b", "x": "1"}, "T": {"x": "2"}}
    assert pc.ini_get(d, "S", "missing") == "default_value"
    assert pc.split_string("a  b") == ["a", "", "b"]
    assert pc.count_spaces("Alpha Beta") == 1 and pc.count_words("  v  w ") == 2 and pc.count_words("") == 0
    assert pc.to_floats(["1.5", "x", "2e-3junk"]) == [np.float32(1.5), np.float32(2e-3)]


@pytest.mark.parametrize("name,N,steps,simindex,scheme", [
    ("HeatEq.ini", 12 * 12 * 33, 1, 0, (12, 2, 40, 0)),
    ("AdvEq.ini", 33 ** 3, 1, 1, (11, 1, 27, 0)),
    ("LUDecomp.ini", 125, 1, 0, (12, 3, 27, 2)),
])
def test_reference_cases_parse(name, N, steps, simindex, scheme):
    spec = pc.setup(CASES, name)
    assert spec.N == N and spec.simindex == simindex
    assert (spec.time_scheme, spec.space_scheme, spec.solver_scheme, spec.data_type) == scheme
    assert spec.dt == np.float32(0.001) and spec.save == 1
    # [Simulation]: the constant Alpha = 0.1 first, then the scalars, then vectors
    assert spec.slots[0].name == "Alpha" and np.all(spec.slots[0].device == np.float32(0.1))
    assert spec.has_bc  # setup.ini lists six BC types
    # float time loop: TotalTime = Timestep gives one step
    t, k = np.float32(0), 0
    while t < spec.totaltime:
        t, k = np.float32(t + spec.dt), k + 1
    assert k == steps


@pytest.mark.parametrize("n", [(12, 12, 33), (5, 5, 5), (4, 7, 3), (3, 3, 3)])
def test_face_lists(n):
    lists = pc.face_lists(n)
    # concatenated in list order they are the C oracle's flat face list
    assert np.array_equal(np.concatenate(lists), po.face_indices(*n))
    nx, ny, nz = n
    assert [len(f) for f in lists] == [(ny - 2) * (nz - 2)] * 2 + [(nx - 2) * (nz - 2)] * 2 + [(nx - 2) * (ny - 2)] * 2
    x, y, z = lists[5] % nx, (lists[5] // nx) % ny, lists[5] // (nx * ny)
    assert np.all(z == nz - 1) and np.all((x > 0) & (x < nx - 1) & (y > 0) & (y < ny - 1))


def test_heateq_bc_snapshot_from_setup_ini():
    spec = pc.setup(CASES, "HeatEq.ini")
    # setup.ini: property T on all faces, value 1.0 on face 5 (z = nz-1)
    faces = pc.face_lists(spec.n)
    snap = spec.snapshot
    assert len(faces[5]) == 100 and np.all(snap[faces[5]] == 1.0)
    assert np.count_nonzero(snap) == 100
    # the device field holds the IC without the BC faces (uniform vector x
    # component: direction 0 0 -1 -> 0)
    assert np.all(spec.slots[1].device == 0.0)


def test_adveq_fields():
    spec = pc.setup(CASES, "AdvEq.ini")
    N = spec.N
    u, v = spec.slots[1], spec.slots[2]
    assert (u.name, u.type, v.name, v.type) == ("U", 1, "v", 2)
    # Gaussian IC (Sigma 5, Median 15, Scalefactor 1000) on l = 31, padded n = 33
    ref = (np.float32(1000) * po.ic_gaussian((33, 33, 33), np.float32([31] * 3), np.float32([15] * 3),
                                             np.float32([5] * 3), 1.0)).astype(np.float32)
    assert u.device.tobytes() == ref.tobytes()
    assert np.all(v.device[:2 * N] == 0) and np.all(v.device[2 * N:] == np.float32(-100))
    # setup.ini's BC property T is not a slot of this case: the face values go
    # to slot 0 and the snapshot is U's initial condition
    assert spec.snapshot.tobytes() == u.device.tobytes()


def _continuity_assets(tmp_path):
    """AdvEq.ini turned into the continuity equation (simindex 2): scalar rho
    (Gaussian IC), vector u (uniform (0,0,-100)), BCs on rho"""
    import re
    import shutil
    d = tmp_path / "assets"
    shutil.copytree(CASES, d, copy_function=shutil.copyfile)
    p = d / "AdvEq.ini"
    t = p.read_text()
    for k, v in (("simindex", "2"), ("Scalars", "rho"), ("Vectors", "u"), ("TotalTime", "0.003")):
        t, n = re.subn(rf"(?m)^{k}\s*=.*$", f"{k} = {v}", t)
        assert n == 1, k
    p.write_text(t)
    s = d / "setup.ini"
    t, n = re.subn(r"(?m)^property\s*=.*$", "property = rho rho rho rho rho rho", s.read_text())
    assert n == 1
    s.write_text(t)
    return str(d)


def test_continuity_operator_equals_advection_with_v_u(tmp_path):
    # process::div("rho", "u") scales each divergence matrix by u's component
    # (SpMSCSR) and merges without factors (csrgeam_2); the advection path
    # merges the same gradient stencils with factors v (csrgeam_3).  With the
    # same norm and v = u the two routes must give the same matrix bit for bit
    spec = pc.setup(_continuity_assets(tmp_path), "AdvEq.ini")
    assert spec.simindex == 2
    kind, D = pc._operator(spec, True)
    n = spec.n
    g = po.grad_norm(spec.l[0], n[0] - 2)
    G = [po.grad_csr(*n, spec.space_scheme, d, g) for d in range(3)]
    A = po.geam3(G[0], G[1], G[2], spec.slots[spec.match("u")].device)
    np.testing.assert_array_equal(D.row_ptr, A.row_ptr)
    np.testing.assert_array_equal(D.col, A.col)
    assert D.val.tobytes() == A.val.tobytes()
    # u = (0, 0, -100): only the z stencil survives with non-zero values
    assert np.count_nonzero(D.val) == np.count_nonzero(G[2].val)


def test_continuity_case_runs(tmp_path):
    spec = pc.setup(_continuity_assets(tmp_path), "AdvEq.ini")
    steps, x, exports = pc.run(spec)
    assert steps == 3 and np.all(np.isfinite(x))
    assert len(exports) == 3


def coulomb_cases():
    with open(os.path.join(GOLD, "ref_digests.json")) as f:
        return json.load(f)["coulomb"]


@pytest.mark.parametrize("meta", coulomb_cases(), ids=lambda m: m["key"])
def test_coulomb_ic_matches_reference(meta):
    # the Coulomb initial condition (src/ic.cpp:29-32, 62-79) against PIRO's
    # own calculatecoulomb, compiled from its sources (tests/golden fixture),
    # with Z, e and epsilon_0 read as distribution.ini text
    ref = np.load(os.path.join(GOLD, "ref_assembly.npz"))[meta["key"]]
    ic = {"Coulomb": {"center": " ".join(repr(float(c)) for c in meta["center"]), "Z": meta["Z"],
                      "e": meta["e"], "epsilon_0": meta["epsilon_0"]}}
    v = pc.initial_condition("Coulomb", tuple(meta["dims"]), tuple(meta["l"]), ic)
    assert v.tobytes() == ref.tobytes()
