"""Freeze the reference's own host-side assembly output as fixtures.

Runs oracle/_ref/piro_ref_harness (built from the PIRO sources by
``make -C oracle ref``) on small grids and stores what it printed:

* ``ref_assembly.npz``  -- full arrays for the small cases;
* ``ref_digests.json``  -- sha256 of the raw arrays for the larger cases
  (32^3 CSR Laplacian of every scheme), so the 2 MB arrays need not be
  committed.

Reference functions exercised (paths relative to the PIRO tree):
  CSR::stencils::laplacian_{cd,uw,dw}   src/stencils.cpp:233-429
  CSR::stencils::gradient_*             src/stencils.cpp:7-231
  DENSE::stencils::laplacian_*          src/stencils.cpp:431-571
  COO::stencils::laplacian_*            src/stencils.cpp:572-795
  HT::stencils::laplacian_* + sethash   src/stencils.cpp:797-993,
                                        src/operatorlaplacian.cpp:194-204
  methods::sethash                      src/methods.cpp:74-118
  calculategaussian                     src/ic.cpp:16-27
  calculatecoulomb                      src/ic.cpp:29-32, 62-79
  bc::prepbc + paraview::writevth/vti   src/bc.cpp:80-105, src/paraview.cpp:12-137

* ``ref_vtk/<case>/``    -- the reference writer's .vth/.vti files for small
  fields, with the fields themselves (``fields.npz``).

Only this script runs the harness; tests read the frozen files.
"""
from __future__ import annotations

import hashlib
import json
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "piro_ref_harness")

LAP_GRIDS = [(3, 3, 3), (4, 4, 4), (5, 6, 7), (7, 5, 4), (9, 9, 9), (6, 3, 5)]
GRAD_GRIDS = [(5, 6, 7), (4, 4, 4)]
BIG_LAP = [(32, 32, 32)]
HT_CASES = [  # (scheme, padded dims, l0): Test/LUDecomp.ini is n=3 3 3, l=5
    (3, (5, 5, 5), 5.0),
    (1, (5, 5, 5), 5.0),
    (2, (6, 5, 4), 4.0),
]
# COO and HT stencil functions of the reference (their own copies of the
# 7-point rows and the boundary 10.0 rule); HT tables sized as HT::laplacian
COO_GRIDS = [(3, 3, 3), (5, 6, 7), (7, 5, 4), (6, 3, 5)]
HTST_GRIDS = [(3, 3, 3), (5, 5, 5), (6, 5, 4), (4, 7, 3)]
# (padded dims, l, centre, Z, e, epsilon_0): the first is the reference's
# distribution.ini [Coulomb] section on a 30-long box; the second puts the
# centre on a grid coordinate (r = 0, where the 1e-10 offset matters)
COULOMB_CASES = [
    ((8, 9, 10), (30.0, 30.0, 30.0), (15.0, 15.0, 15.0), "1", "8.854187817E-12", "1.602176634E-19"),
    ((12, 7, 9), (1.0, 1.4, 0.9), (0.5, 0.6, 0.3), "2", "1.602176634E-19", "8.854187817E-12"),
]
GAUSS_CASES = [
    ((8, 9, 10), (1.0, 1.2, 0.8), (0.5, 0.55, 0.4), (0.1, 0.2, 0.15)),
    ((32, 32, 32), (30.0 / 31.0,) * 3, (15.0 / 31.0,) * 3, (0.1,) * 3),
]


# (padded dims, scalars, vectors, step): interior values include signed
# zeros, infinities, NaNs of both signs, subnormals and wide exponents
VTK_CASES = [
    ("mixed_5x4x6", (5, 4, 6), 2, 1, 7),
    ("single_3x3x3", (3, 3, 3), 1, 0, 12345),
    ("vectors_4x5x3", (4, 5, 3), 1, 2, 0),
]


def vtk_fields(dims, n_scal, n_vec, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    N = nx * ny * nz
    interior = [x + y * nx + z * nx * ny for z in range(1, nz - 1) for y in range(1, ny - 1) for x in range(1, nx - 1)]
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, -np.nan, 1e-40, -3e-39, 1.0 / 3, 123456789.0,
                        1e30, -2.5e-12, 100000.0, 999999.0, 1e6, 0.1], np.float32)
    out = []
    for k in range(n_scal + n_vec):
        m = 3 * N if k >= n_scal else N
        v = (rng.standard_normal(m) * 10.0 ** rng.integers(-9, 10, m)).astype(np.float32)
        for j, idx in enumerate(interior[: len(special)]):
            v[idx + (N if k >= n_scal else 0)] = special[(j + k) % len(special)]
        out.append(v)
    return out


# DENSE Laplacian grids (N x N floats; frozen as non-zero triplets + digest)
DENSE_GRIDS = [(4, 5, 6), (6, 6, 6)]


def f32_hex(x: float) -> str:
    return "%08x" % struct.unpack("<I", struct.pack("<f", np.float32(x)))[0]


def lap_norm(l0: float, n0: int) -> np.float32:
    # operatorlaplacian.cpp:33
    h = np.float32(l0) / np.float32(n0)
    return np.float32(float(h) ** 2)


def grad_norm(l0: float, n0: int) -> np.float32:
    # operatorgradient.cpp:25
    return np.float32(2.0) * (np.float32(l0) / np.float32(n0))


def ht_table_size(sparsecount: int, N: int) -> int:
    # operatorlaplacian.cpp:184-185 with assets/Data/hashtable.ini
    a, b, c, lf = np.float32(3.3608), np.float32(-0.5552), np.float32(-0.0143), np.float32(0.7)
    load = np.float32(float(a) * (float(N) ** float(b)) + float(c))
    return int(np.float32(sparsecount) / np.float32(load * lf))


def read_arrays(path: str, dtypes):
    out = []
    with open(path, "rb") as f:
        for dt in dtypes:
            (n,) = struct.unpack("<q", f.read(8))
            out.append(np.frombuffer(f.read(n * np.dtype(dt).itemsize), dtype=dt).copy())
    return out


def run(args, dtypes):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "o.bin")
        subprocess.run([HARNESS] + [str(a) for a in args] + [out], check=True)
        return read_arrays(out, dtypes)


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main() -> int:
    if not os.path.exists(HARNESS):
        print("build the harness first: make -C oracle ref", file=sys.stderr)
        return 1
    arrays = {}
    meta = {"lap": [], "grad": [], "ht": [], "gauss": [], "lap_digest": []}
    for scheme in (1, 2, 3):
        for (nx, ny, nz) in LAP_GRIDS:
            norm = lap_norm(1.0, nx - 2)
            rp, col, val = run(["lap", scheme, nx, ny, nz, f32_hex(norm)], ["<i4", "<i4", "<f4"])
            key = f"lap_s{scheme}_{nx}x{ny}x{nz}"
            arrays[key + "_rp"], arrays[key + "_col"], arrays[key + "_val"] = rp, col, val
            meta["lap"].append({"key": key, "scheme": scheme, "dims": [nx, ny, nz],
                                "norm_hex": f32_hex(norm)})
        for (nx, ny, nz) in BIG_LAP:
            norm = lap_norm(30.0 / 31.0, nx - 2)
            rp, col, val = run(["lap", scheme, nx, ny, nz, f32_hex(norm)], ["<i4", "<i4", "<f4"])
            meta["lap_digest"].append({"scheme": scheme, "dims": [nx, ny, nz], "norm_hex": f32_hex(norm),
                                       "rp": digest(rp), "col": digest(col), "val": digest(val),
                                       "nnz": int(col.size)})
        for d in range(3):
            for (nx, ny, nz) in GRAD_GRIDS:
                norm = grad_norm(1.0, nx - 2)
                rp, col, val = run(["grad", scheme, d, nx, ny, nz, f32_hex(norm)], ["<i4", "<i4", "<f4"])
                key = f"grad_s{scheme}_d{d}_{nx}x{ny}x{nz}"
                arrays[key + "_rp"], arrays[key + "_col"], arrays[key + "_val"] = rp, col, val
                meta["grad"].append({"key": key, "scheme": scheme, "dir": d, "dims": [nx, ny, nz],
                                     "norm_hex": f32_hex(norm)})
    for scheme in (1, 2, 3):
        for (nx, ny, nz) in DENSE_GRIDS:
            norm = lap_norm(1.0, nx - 2)
            (m,) = run(["dense", scheme, nx, ny, nz, f32_hex(norm)], ["<f4"])
            N = nx * ny * nz
            m = m.reshape(N, N)
            r, c = np.nonzero(m)
            key = f"dense_s{scheme}_{nx}x{ny}x{nz}"
            arrays[key + "_r"], arrays[key + "_c"], arrays[key + "_v"] = r.astype(np.int32), c.astype(np.int32), m[r, c]
            meta.setdefault("dense", []).append({"key": key, "scheme": scheme, "dims": [nx, ny, nz],
                                                 "norm_hex": f32_hex(norm), "digest": digest(m)})
    for scheme, (nx, ny, nz), l0 in HT_CASES:
        norm = lap_norm(l0, nx - 2)
        N = nx * ny * nz
        nnz = 7 * N - 2 * (ny * nz + nx * nz + nx * ny)
        ts = ht_table_size(nnz, N)
        keys, vals = run(["ht", scheme, nx, ny, nz, f32_hex(norm), ts], ["<i4", "<f4"])
        key = f"ht_s{scheme}_{nx}x{ny}x{nz}"
        arrays[key + "_keys"], arrays[key + "_vals"] = keys, vals
        meta["ht"].append({"key": key, "scheme": scheme, "dims": [nx, ny, nz], "norm_hex": f32_hex(norm),
                           "table_size": ts, "sparsecount": nnz, "l0": l0})
    for scheme in (1, 2, 3):
        for (nx, ny, nz) in COO_GRIDS:
            norm = lap_norm(1.0, nx - 2)
            r, c, v = run(["coo", scheme, nx, ny, nz, f32_hex(norm)], ["<i4", "<i4", "<f4"])
            key = f"coo_s{scheme}_{nx}x{ny}x{nz}"
            arrays[key + "_rows"], arrays[key + "_cols"], arrays[key + "_vals"] = r, c, v
            meta.setdefault("coo", []).append({"key": key, "scheme": scheme, "dims": [nx, ny, nz],
                                               "norm_hex": f32_hex(norm)})
        for (nx, ny, nz) in HTST_GRIDS:
            norm = lap_norm(1.0, nx - 2)
            N = nx * ny * nz
            nnz = 7 * N - 2 * (ny * nz + nx * nz + nx * ny)
            ts = ht_table_size(nnz, N)
            rp, c, v, keys, vals = run(["htst", scheme, nx, ny, nz, f32_hex(norm), ts],
                                       ["<i4", "<i4", "<f4", "<i4", "<f4"])
            key = f"htst_s{scheme}_{nx}x{ny}x{nz}"
            for name, a in (("_rp", rp), ("_col", c), ("_val", v), ("_keys", keys), ("_vals", vals)):
                arrays[key + name] = a
            meta.setdefault("htst", []).append({"key": key, "scheme": scheme, "dims": [nx, ny, nz],
                                                "norm_hex": f32_hex(norm), "table_size": ts})
    for i, (dims, l, centre, Z, e, eps0) in enumerate(COULOMB_CASES):
        (g,) = run(["coulomb", *dims, *l, *centre, Z, e, eps0], ["<f4"])
        key = f"coulomb_{i}"
        arrays[key] = g
        meta.setdefault("coulomb", []).append({"key": key, "dims": list(dims), "l": list(l),
                                               "center": list(centre), "Z": Z, "e": e, "epsilon_0": eps0})
    for i, (dims, l, mean, sigma) in enumerate(GAUSS_CASES):
        (g,) = run(["gauss", *dims, *l, *mean, *sigma], ["<f4"])
        key = f"gauss_{i}"
        if g.size <= 4096:
            arrays[key] = g
        meta["gauss"].append({"key": key, "dims": list(dims), "l": list(l), "mean": list(mean),
                              "sigma": list(sigma), "digest": digest(g),
                              "sum_f64": float(np.sum(g.astype(np.float64)))})
    for i, (key, dims, n_scal, n_vec, step) in enumerate(VTK_CASES):
        fields = vtk_fields(dims, n_scal, n_vec, 500 + i)
        outdir = os.path.join(HERE, "ref_vtk", key)
        os.makedirs(outdir, exist_ok=True)
        with tempfile.TemporaryDirectory() as td:
            inp = os.path.join(td, "in.bin")
            np.concatenate(fields).astype("<f4").tofile(inp)
            subprocess.run([HARNESS, "vtk", *map(str, dims), str(n_scal), str(n_vec), str(step), inp, td],
                           check=True, capture_output=True)
            tag = "%05d" % step
            for src, dst in ((f"case/mesh/mesh_{tag}.vth", f"mesh_{tag}.vth"),
                             (f"case/mesh/level/{tag}_level_0.vti", f"{tag}_level_0.vti")):
                with open(os.path.join(td, src), "rb") as fi, open(os.path.join(outdir, dst), "wb") as fo:
                    fo.write(fi.read())
        np.savez_compressed(os.path.join(outdir, "fields.npz"),
                            **{f"f{k}": f for k, f in enumerate(fields)})
        meta.setdefault("vtk", []).append({"key": key, "dims": list(dims), "n_scal": n_scal, "n_vec": n_vec,
                                           "step": step, "names": ["Alpha"] + [f"S{k}" for k in range(n_scal)]
                                           + [f"V{k}" for k in range(n_vec)]})
    np.savez_compressed(os.path.join(HERE, "ref_assembly.npz"), **arrays)
    with open(os.path.join(HERE, "ref_digests.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print(f"wrote {len(arrays)} arrays")
    return 0


if __name__ == "__main__":
    sys.exit(main())
