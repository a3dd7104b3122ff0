"""The Fortran drop-in: a Fortran program calling lfmm3dparttarg and
l3dpartdirect with the reference's calling sequences, linked against
libfmm3d_mi355x_f.so (ISO_C_BINDING over the C ABI)."""
import os
import subprocess

import pytest

import fmmlib3d_amd.build as B

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.exists(B.FDRIVER), reason="Fortran driver not built (no Fortran compiler)")
def test_fortran_driver_runs_on_gpu():
    out = subprocess.run([B.FDRIVER], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert out.returncode == 0, out.stdout
    line = [l for l in out.stdout.splitlines() if l.startswith("relerr")][-1]
    errs = [float(x) for x in line.split()[-4:]]
    # iprec = 1: tolerance 0.5e-3 for the potential, a digit less for the field
    assert errs[0] < 5e-4 and errs[2] < 5e-4
    assert errs[1] < 5e-3 and errs[3] < 5e-3
    line = [l for l in out.stdout.splitlines() if l.startswith("hrelerr")][-1]
    errs = [float(x) for x in line.split()[-2:]]
    assert errs[0] < 5e-4 and errs[1] < 5e-3
