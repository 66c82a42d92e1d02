"""CPU checks of the arithmetic shortcuts in the lane sweeps' Godunov update
(k_fsm.hip godunov_cand), which must leave every result bit-identical to the
oracle's plain IEEE form (cuda_rasterizer.cu:405-425).

The division by 3 of the 3-D candidate is fma(x, zh, x * zl) with zh = RN(1/3)
and zl = RN(1/3 - zh).  tests/cxx/div3_check.c compares it with IEEE division
over every non-negative finite binary32 value; the sweeps only select a 3-D
numerator >= h / 2 >= 2^-31 (lattices with 2^-30 <= h take those kernels).
"""
import os
import shutil
import struct
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cxx", "div3_check.c")


def _has_fma() -> bool:
    try:
        with open("/proc/cpuinfo") as fh:
            return any(line.startswith("flags") and " fma " in f" {line.split(':', 1)[1]} " for line in fh)
    except OSError:
        return False


@pytest.fixture(scope="module")
def div3_check(tmp_path_factory):
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not available")
    out = str(tmp_path_factory.mktemp("div3") / "div3_check")
    flags = ["-mfma"] if _has_fma() else []
    subprocess.run([cc, "-O2", "-std=c11", "-ffp-contract=off", *flags, "-pthread", "-o", out, SRC, "-lm"],
                   check=True)
    return out


def _run(binary, zl_hex, stride=1):
    r = subprocess.run([binary, zl_hex, str(stride)], check=True, capture_output=True, text=True, timeout=600)
    bad, last = r.stdout.split()
    return int(bad), struct.unpack("<f", struct.pack("<I", int(last)))[0]


def test_div3_split_equals_ieee_division_above_2m123(div3_check):
    bad, last = _run(div3_check, "-0x1.555556p-27")
    # every mismatch lies where x * zl underflows, far below the sweeps' range
    assert bad > 0 and last <= 2.0 ** -123, (bad, last)


def test_div3_check_catches_a_wrong_low_part(div3_check):
    # the check has teeth: a plain multiply by RN(1/3) (zl dropped), a low part
    # of the wrong sign, and one off by a factor of two give wrong quotients
    # well inside the sweeps' range (a sample of every 1009th bit pattern)
    for zl in ("0x0p+0", "0x1.555556p-27", "-0x1.555556p-26", "-0x1.555556p-28"):
        bad, last = _run(div3_check, zl, 1009)
        assert last > 2.0 ** -30, (zl, bad, last)
