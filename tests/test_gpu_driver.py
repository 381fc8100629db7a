"""The C++ driver (the counterpart of RR27V33_code_DNS_LES/main.cu) end to
end: the fused path and the drop-in kernels in the main.cu order must print
the same diagnostics (FP64 storage), for each wall treatment and operator."""
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(PKG, "rrlbm_channel")


def run_driver(tmp_path, name, *extra):
    out = tmp_path / name
    cmd = [DRIVER, "--lx", "32", "--ly", "16", "--lz", "8", "--steps", "20", "--ndiag", "10", "--nout", "20",
           "--storage", "f64", "--init", "channel", "--dir", str(out)] + list(extra)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    stats = np.loadtxt(out / "statistics.dat")
    assert stats.shape == (2, 5)
    assert (out / ("%09d.dat" % 20)).exists() and (out / ("f%09d.dat" % 20)).exists()
    return stats, r.stdout


@pytest.mark.parametrize("variant", [[], ["--wall", "fdbc"], ["--rr-order", "6"], ["--bgk"], ["--les"],
                                     ["--gp-steps", "15"], ["--gp-steps", "15", "--les"]])
def test_driver_fused_equals_reference_order(tmp_path, variant):
    if not os.path.exists(DRIVER):
        pytest.fail("driver not built: run the build first")
    fused, out = run_driver(tmp_path, "fused", *variant)
    ref, _ = run_driver(tmp_path, "ref", "--reference", *variant)
    assert "umean" in out
    # fused LES (the default lagged gradient) differs by the one-step lag of tau_t
    rtol, atol = (1e-4, 1e-8) if "--les" in variant else (1e-12, 1e-15)
    assert np.allclose(fused[:, 1:], ref[:, 1:], rtol=rtol, atol=atol), (fused, ref)
