"""The C++ ABI client (tests/abi_client/resident_fit.cpp): g++ against
include/bonsai_hip.h, linked to libbonsai_hip.so, driving fits through an adapter shaped
like the reference's grower seam (resident_begin / grow with empty gradients /
resident_end, include/bonsai/grower.hpp:254-306 and include/bonsai/booster.hpp:630-676),
every tree checked against the oracle inside the program."""
import subprocess
from pathlib import Path

import pytest

CLIENT = Path(__file__).resolve().parent / "abi_client" / "resident_fit"


def run(*args):
    assert CLIENT.exists(), "build first (__graft_entry__.build())"
    return subprocess.run([str(CLIENT), *args], capture_output=True, text=True, timeout=300)


def test_client_links_and_reports_no_device_on_the_cpu():
    """On a machine without a GPU the program loads both libraries and says so (77)."""
    import bonsai_amd
    if bonsai_amd.device_count() > 0:
        pytest.skip("a device is visible: the gpu test runs the fit")
    r = run()
    assert r.returncode == 77, r.stdout + r.stderr
    assert "no HIP device" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("grower", ["depthwise", "leafwise"])
def test_client_resident_fit_matches_the_oracle(grower):
    print(f"[abi client] {grower}", flush=True)
    r = run(grower)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "match the oracle" in r.stdout
