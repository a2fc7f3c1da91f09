"""The C++ mirror of sph2d_box (include/emerald_sph2d_box.hpp) as the
reference's callers use it: tests/cpp/api_check runs EZ_EXAMPLE_SIM for one
frame with a host User_forces_function lambda (gravity by hand) and a host
User_colors_function, and again with the default device gravity. Both final
states equal the Python mirror's EZ_EXAMPLE_SIM bit for bit (the callback path
integrates the same force as the fused device gravity), the callback ran once
per sub-step, and the colours are the callback's."""

import os
import subprocess

import numpy as np
import pytest

from emerald_amd import sph2d_box as sb
from helpers import assert_bits_equal

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "api_check")


def test_cpp_mirror_callbacks(gpu, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/api_check is not built (python -m emerald_amd.build)")
    a = str(tmp_path / "host.emsph")
    b = str(tmp_path / "device.emsph")
    print("[cpp api] api_check", flush=True)
    r = subprocess.run([EXE, a, b], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "colors ok" in r.stdout
    sa, ta, _ = sb.load_state(a)
    sd, td, _ = sb.load_state(b)
    sim = sb.EZ_EXAMPLE_SIM(sb.Parameters(method="iisph"))
    sim.step()
    st = sim.fetch_state()
    for s, what in ((sa, "host callback"), (sd, "device gravity")):
        assert_bits_equal(s.positions, st.positions, f"C++ {what} positions")
        assert_bits_equal(s.velocities, st.velocities, f"C++ {what} velocities")
    assert_bits_equal(sa.colors[:, 0], sa.velocities[:, 0], "callback colours")
    assert np.all(sa.colors[:, 2] == np.float32(0.25))
    assert_bits_equal(sd.colors, st.colors, "default colours")
