"""The ISO_C_BINDING interface module (adcirc_amd/fortran/adcgpu_iface.F90)
driving the drop-in JCG from Fortran, as src/gwce.F:2002-2005 calls ITPACK."""
import os
import subprocess

import pytest

from gpu_helpers import log

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "adcirc_amd", "fortran", "jcg_dropin_check")


@pytest.mark.gpu
def test_fortran_interface_drives_dropin_jcg():
    if not os.path.exists(EXE):
        pytest.skip("Fortran check not built (no flang at build time)")
    log("case fortran: adcgpu_jcg through the bind(C) interface, 1000-row system")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    log("  " + r.stdout.strip().splitlines()[0] if r.stdout.strip() else "  (no output)")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ier=0" in r.stdout
