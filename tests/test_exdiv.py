"""The FMA-corrected division of the HIP kernels (exdiv) equals IEEE division bit for bit:
swept on the CPU with the same three-operation formula (tests/exdiv_check.c)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_exdiv_equals_ieee_division(tmp_path):
    exe = str(tmp_path / "exdiv_check")
    try:
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math",
                        os.path.join(HERE, "exdiv_check.c"), "-lm", "-o", exe], check=True)
    except (OSError, subprocess.CalledProcessError) as e:
        pytest.skip(f"no C compiler: {e}")
    out = subprocess.run([exe, "20000000"], capture_output=True, text=True, check=True).stdout
    assert out.strip().splitlines()[-1] == "0", out
