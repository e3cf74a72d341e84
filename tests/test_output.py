"""Output snapshots (include/shark_output.h) against the reference's own `output(iout)`
(src/commons/write_output.f90:2-300).

The fixtures tests/golden/ref_output_<case>.npz hold the files the flang-built reference wrote for
the final state of tests/golden/ref_<case>.npz.  shark_output_write, given the primitives of that
state (ctoprim, solve.f90:195-253, restated here with the same IEEE operations), must write the same
files byte for byte.  The info.dat numbers use Fortran's E23.15, checked separately against flang on
values chosen to exercise rounding, three-digit exponents, zero and negative numbers."""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from refcases import GOLDEN, load_ref
from shark_amd import gpu

OUT_CASES = sorted(f[11:-4] for f in os.listdir(GOLDEN) if f.startswith("ref_output_") and f.endswith(".npz"))


def ctoprim(u, cs, nd, smallr, cfg=None):
    """q from u_prim as the reference's ctoprim (solve.f90:210-239): isothermal, or with
    iso_cs < 1 the pressure from the total energy."""
    q = np.empty_like(u)
    rho = u[0]
    q[0] = np.maximum(rho, smallr)
    q[1], q[2], q[3] = u[1] / rho, u[2] / rho, u[3] / rho
    if cfg is not None and cfg.iso_cs < 1:
        ekin = 0.5 * rho * (q[1] * q[1]) + 0.5 * rho * (q[2] * q[2]) + 0.5 * rho * (q[3] * q[3])
        q[4] = np.maximum((cfg.gamma - 1.0) * (u[4] - ekin - 0.0), cfg.smallp)
    else:
        q[4] = rho * (cs * cs)
    for k in range(nd):
        rd = u[5 + k]
        q[5 + k] = rd
        for c in range(1, 4):
            q[5 + c * nd + k] = u[5 + c * nd + k] / rd
    return q


def load_expected(case):
    z = np.load(os.path.join(GOLDEN, f"ref_output_{case}.npz"))
    m = json.loads(str(z["meta"]))
    files = {f: bytes(z["file_" + f.replace(".", "_")]) for f in m["files"]}
    return m, files


def compare_dir(d, files):
    got = sorted(os.listdir(d))
    assert got == sorted(files), (got, sorted(files))
    for f, ref in files.items():
        data = open(os.path.join(d, f), "rb").read()
        if f == "info.dat":
            assert data.decode() == ref.decode()
        else:
            a, b = np.frombuffer(data, np.float64), np.frombuffer(ref, np.float64)
            assert a.shape == b.shape, f
            bad = np.flatnonzero(a.view(np.uint64) != b.view(np.uint64))
            assert bad.size == 0, (f, bad[:5], a[bad[:5]], b[bad[:5]])


@pytest.mark.parametrize("case", OUT_CASES)
def test_output_files_match_reference(case):
    cfg, geom, state, final, _, meta = load_ref(case)
    m, files = load_expected(case)
    q = ctoprim(final, state["cs"], cfg.ndust, cfg.smallr, cfg)
    info = gpu.output_info(m["time"], m["units"], m["gridspace"])
    with tempfile.TemporaryDirectory() as td:
        info.base_dir = os.fsencode(td)
        gpu.output_write(info, m["iout"], cfg, q, state["sdust"], geom["position"], geom["radii_c"])
        compare_dir(os.path.join(td, f"output_{m['iout']:05d}"), files)


def test_units_of_reference_setups():
    """The unit sets the Python host offers equal the reference's units modules bit for bit."""
    for case, setup in (("disk_2d", "disk_polar"), ("si_2d", "SI"), ("dustywave_dhll", "DUSTYWAVE"),
                        ("sod_hll_s3", "SOD"), ("kh_n1_hllc", "KH"), ("cdisk_visc_n1", "disk")):
        m, _ = load_expected(case)
        assert gpu.units_of(setup) == m["units"], case


def test_output_rejects_unbuilt_fields():
    cfg, geom, state, final, _, _ = load_ref("dustywave_dhll")
    info = gpu.output_info(0.0, gpu.units_of("DUSTYWAVE"))
    info.charging = 1
    with tempfile.TemporaryDirectory() as td:
        info.base_dir = os.fsencode(td)
        with pytest.raises(gpu.SharkGPUError, match="charging"):
            gpu.output_write(info, 1, cfg, final, state["sdust"], geom["position"], geom["radii_c"])
        info.charging = 0
        with pytest.raises(gpu.SharkGPUError, match="iout"):
            gpu.output_write(info, 100000, cfg, final, state["sdust"], geom["position"], geom["radii_c"])


VALUES = [0.0, -0.0, 1.0, -1.0, 0.1, 2.0 / 3.0, 9.9999999999999995e-8, 0.99999999999999994,
          9.9999999999999967, 123456.78901234567, 5.925925925925926e-07, 8893333333333.332,
          1.5e13, 1e-99, 1e-100, 3.3e100, -7.25e-150, 1.7976931348623157e308, 4.9e-324,
          0.5e-15, 1.0000000000000005]


@pytest.mark.skipif(shutil.which("flang") is None, reason="flang (the reference's compiler) is not installed")
def test_e23_15_formatting_matches_flang():
    """info.dat's numbers (time, units) in Fortran E23.15 as flang prints them."""
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "fmt.f90")
        lines = ["program fmt", "  implicit none", "  real(8) :: v(%d)" % len(VALUES), "  integer :: i"]
        for i, v in enumerate(VALUES):   # exact bit patterns
            lines.append("  v(%d) = transfer(int(z'%016X', 8), 1.0d0)" % (i + 1, np.float64(v).view(np.uint64)))
        lines += ["  do i = 1, size(v)", "    write(*,'(\"time        =\",E23.15)') v(i)",
                  "  end do", "end program fmt"]
        open(src, "w").write("\n".join(lines) + "\n")
        subprocess.run(["flang", "-o", os.path.join(td, "fmt"), src], check=True, capture_output=True)
        want = subprocess.run([os.path.join(td, "fmt")], check=True, capture_output=True,
                              text=True).stdout.splitlines()
        assert len(want) == len(VALUES)
        cfg, geom, state, final, _, _ = load_ref("dustywave_dhll")
        for v, w in zip(VALUES, want):
            info = gpu.output_info(v, gpu.units_of("DUSTYWAVE"), base_dir=td)
            gpu.output_write(info, 1, cfg, final, state["sdust"], geom["position"], geom["radii_c"])
            line = [l for l in open(os.path.join(td, "output_00001", "info.dat")) if l.startswith("time")][0]
            assert line.rstrip("\n") == w, (v, line, w)
