"""GPU parity of the case driver (piro_case_run): the reference's own test
cases (tests/golden/cases, from Test/*.ini and dependencies/assets) and
variants of them that run more steps or other schemes, against the CPU
restatement of PIRO's preprocess + time loop (oracle/case.py).

Bar: forward-Euler runs whose initial field the host writes exactly (the
uniform-vector IC) are bit-identical; runs from the Gaussian IC (device expf
within a few ulp of the host's) and JLI runs (FP64 device residual, the stop
may fall one iteration apart) agree to 1e-5 relative L2; the LU run (FP32
factors) to 1e-5 against the FP64 direct solve.  The last Paraview file of
every run equals the oracle writer's text for the device's own final field.
"""
import os
import re
import shutil

import numpy as np
import pytest

import oracle as po
from oracle import case as pc

pytestmark = pytest.mark.gpu

CASES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases")


def make_assets(tmp_path, case, edits=None, setup_edits=None):
    """copy the fixture assets and apply `key = value` edits to the case file
    (and to setup.ini)"""
    d = tmp_path / "assets"
    # copyfile: new files with default permissions (the fixtures may be read-only)
    shutil.copytree(CASES, d, copy_function=shutil.copyfile)
    for fname, ed in ((case, edits), ("setup.ini", setup_edits)):
        if not ed:
            continue
        p = d / fname
        text = p.read_text()
        for key, value in ed.items():
            text, k = re.subn(rf"(?m)^{re.escape(key)}\s*=.*$", f"{key} = {value}", text)
            assert k == 1, key
        p.write_text(text)
    return str(d)


def run_both(tmp_path, case, edits=None, setup_edits=None, alpha_row_scaled=True):
    from piro_amd.core import case_run
    assets = make_assets(tmp_path, case, edits, setup_edits)
    spec = pc.setup(assets, case)
    out = str(tmp_path / "out")
    print(f"case {case} {edits} {setup_edits} alpha_row_scaled={alpha_row_scaled}", flush=True)
    steps, field = case_run(assets, case, out, -1, alpha_row_scaled, spec.N)
    ref_steps, ref, exports = pc.run(spec, -1, alpha_row_scaled)
    assert steps == ref_steps
    check_paraview(out, spec, field, exports)
    return spec, field, ref, steps


def check_paraview(out, spec, field, exports):
    base = os.path.join(out, spec.casename, "mesh")
    files = sorted(os.listdir(os.path.join(base, "level")))
    assert files == sorted(f"{pc.step_tag(ts)}_level_0.vti" for ts, _ in exports)
    ts, printed = exports[-1]
    typed = [(name, t, field if k == 0 else values) for k, (name, t, values) in enumerate(printed)]
    with open(os.path.join(base, "level", f"{pc.step_tag(ts)}_level_0.vti")) as f:
        assert f.read() == pc.writevti(spec.n, spec.origin, spec.spacing, typed) + "\n"
    with open(os.path.join(base, f"mesh_{pc.step_tag(ts)}.vth")) as f:
        assert f.read() == pc.writevth(spec.n, ts) + "\n"


@pytest.mark.parametrize("data", [0, 3, 2])
def test_heat_forward_euler_bit_exact(tmp_path, data):
    # HeatEq.ini switched to forward Euler on CSR (0), COO (3) and HT (2):
    # zero field with the z = nz-1 face held at 1 diffusing for 20 steps
    # (the HT oracle step is the reference's literal O(N^2) table walk: 5 steps)
    total = 0.005 if data == 2 else 0.02
    spec, field, ref, steps = run_both(tmp_path, "HeatEq.ini",
                                       {"Time": 11, "Data": data, "TotalTime": total, "Save": 5})
    assert steps >= 5
    assert field.tobytes() == ref.tobytes()
    assert field.max() == 1.0 and np.count_nonzero(field) > np.count_nonzero(spec.snapshot)


def test_heat_forward_euler_dense_bit_exact(tmp_path):
    # HeatEq.ini on the DENSE format (N = 12*12*33 = 4752, a 90 MB matrix):
    # all rows advance, then the faces are rewritten from the snapshot
    spec, field, ref, steps = run_both(tmp_path, "HeatEq.ini",
                                       {"Time": 11, "Data": 1, "TotalTime": 0.01, "Save": 5})
    assert steps >= 5
    assert field.tobytes() == ref.tobytes()
    assert field.max() == 1.0


def test_continuity_case(tmp_path):
    # AdvEq.ini turned into the continuity equation (simindex 2): rho from the
    # Gaussian IC, u = (0, 0, -100), div("rho", "u") stepped by forward Euler
    spec, field, ref, steps = run_both(
        tmp_path, "AdvEq.ini", {"simindex": 2, "Scalars": "rho", "Vectors": "u", "TotalTime": 0.01, "Save": 2},
        {"property": "rho rho rho rho rho rho"})
    assert steps >= 5
    err = po.rel_l2(field, ref)
    print(f"continuity: rel L2 {err:.2e}")
    assert err <= 1e-5
    assert np.abs(field - ref).max() <= 1e-5 * np.abs(ref).max()


def test_heat_forward_euler_shipped_operator(tmp_path):
    # alpha_row_scaled = 0: the CSR operator without the diffusivity, as the
    # reference's csrscale leaves it
    _, field, ref, _ = run_both(tmp_path, "HeatEq.ini", {"Time": 11, "TotalTime": 0.01}, alpha_row_scaled=False)
    assert field.tobytes() == ref.tobytes()


def test_heat_jli_reference_case(tmp_path):
    # Test/HeatEq.ini as shipped (backward Euler by JLI, one step), then 20 steps
    for edits in (None, {"TotalTime": 0.02, "Save": 4}):
        sub = tmp_path / ("a" if edits is None else "b")
        sub.mkdir()
        _, field, ref, _ = run_both(sub, "HeatEq.ini", edits)
        err = po.rel_l2(field, ref)
        print(f"HeatEq JLI {edits}: rel L2 {err:.2e}")
        assert err <= 1e-5


def test_advection_reference_case(tmp_path):
    # Test/AdvEq.ini: upwind gradients, v = (0, 0, -100) from the IC file,
    # csrgeam_3 concatenation, forward Euler from the Gaussian IC; 1 and 10 steps
    for edits in (None, {"TotalTime": 0.01, "Save": 2}):
        sub = tmp_path / ("a" if edits is None else "b")
        sub.mkdir()
        _, field, ref, _ = run_both(sub, "AdvEq.ini", edits)
        err = po.rel_l2(field, ref)
        print(f"AdvEq {edits}: rel L2 {err:.2e}")
        assert err <= 1e-5
        assert np.abs(field - ref).max() <= 1e-5 * np.abs(ref).max()


def test_lu_reference_case(tmp_path):
    # Test/LUDecomp.ini (HT data, Solver 27): backward Euler by the dense LU
    # over 5 steps, with the top face held at 1
    _, field, ref, _ = run_both(tmp_path, "LUDecomp.ini", {"TotalTime": 0.005})
    err = po.rel_l2(field, ref)
    print(f"LUDecomp: rel L2 {err:.2e}")
    assert err <= 1e-5 and field.max() == 1.0


def test_bc_on_other_slot_is_printed_from_host(tmp_path):
    # setup.ini naming U on two faces of AdvEq: the reference writes those
    # values into U's host copy (slot 1) -- and naming v writes them into the
    # vector's host copy, which the Paraview file prints while the device
    # vector stays as initialised
    run_both(tmp_path, "AdvEq.ini", {"TotalTime": 0.002},
             setup_edits={"property": "U U v v U U", "values": "0.5 0.0 7.0 7.0 0.0 0.0"})


def test_heat_forward_euler_coulomb_ic(tmp_path):
    # the Coulomb IC (the facade's host code, from distribution.ini's
    # [Coulomb] section) stepped by forward Euler: bit-exact, so the device
    # run starts from the same field as the restatement, which the reference's
    # own calculatecoulomb pins (tests/test_oracle_case.py)
    spec, field, ref, steps = run_both(tmp_path, "HeatEq.ini",
                                       {"Time": 11, "filename": "Coulomb", "TotalTime": 0.005, "Save": 5})
    assert steps >= 5
    assert field.tobytes() == ref.tobytes()
    assert np.all(np.isfinite(field)) and field.min() < -1e9
