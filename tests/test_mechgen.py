"""The compiled mechanism header (csrc/sg_mech_h2.h) is the one mechgen.py
writes from data/h2_9sp/chem.dat through the same chemistry fill the ctypes
binding uses, so a configuration built from that file matches it bit for bit."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "senga-versiona_amd")


def test_header_is_current():
    r = subprocess.run([sys.executable, os.path.join(PKG, "mechgen.py"), "--check"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_header_holds_the_case_chemistry():
    sys.path.insert(0, PKG)
    import mechgen
    from sengapy import abi
    from sengapy import case as C
    c = mechgen.config()
    m = C.h2_mechanism()
    assert c.nspec == 9 and c.nstep == 21 and c.nstep == len(m.nsspec)
    # 0-based lists; the Troe and Lindemann constants read through REAL are
    # single-precision values
    assert list(c.nsspec[0][:c.nsslen[0]]) == [s - 1 for s in m.nsspec[0]]
    for j in range(abi.MAX_FALL):
        assert c.rclind[j][3] == float(np.float32(c.rclind[j][3]))
    text = mechgen.render(c)
    assert "struct MechH2" in text and "nstep = 21" in text
