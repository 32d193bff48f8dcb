"""The FP64 division-by-constant sequence of the device PPM code
(div_const in flash5_amd/csrc/fx_device.hpp) against IEEE division, on the
CPU: q = x RN(1/c), rho = fma(-q, c, x), copysign(fma(rho, RN(1/c), q), x)
for c = 3, 6, 12 must give the quotient's exact bits for random dividends
over the whole normal range, for significands of every residue mod 3 at both
ends of many binades, and for signed zeros.  glibc's fma is correctly
rounded, as the device's v_fma_f64 is."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_division_by_constants_matches_ieee(tmp_path):
    exe = str(tmp_path / "cdc")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "cdiv", "const_div_check.c"), "-lm"], check=True)
    out = subprocess.run([exe, "2000000"], capture_output=True, text=True, check=False)
    rows = [line.split() for line in out.stdout.strip().splitlines()]
    assert [r[0] for r in rows] == ["3", "6", "12"]
    for c, tried, bad in rows:
        assert int(tried) > 60_000_000 and int(bad) == 0, (c, tried, bad)
    assert out.returncode == 0
