"""The local Jacobians of the discrete adjoint, on the CPU: each reverse
sweep of libdgx (roe_flux_vjp, face_physics_vjp, volume_flux_vjp,
traction_vjp) against forward-mode dual numbers run through the forward
physics it transposes, on 20000 random states including both entropy-fix
branches of the Roe flux (tests/cpu/vjp_check.cpp, compiled with g++ against
a host-only stand-in for the HIP header).  Bar: 1e-12 of the largest
gradient entry (measured ~2e-15)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_reverse_sweeps_match_dual_numbers(tmp_path):
    exe = str(tmp_path / "vjp_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wno-unknown-pragmas",
                    "-I" + os.path.join(ROOT, "tests", "cpu"),
                    "-I" + os.path.join(ROOT, "dgpr_amd", "csrc"),
                    "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpu", "vjp_check.cpp"), "-o", exe],
                   check=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    for line in r.stdout.split("\n"):
        if line.strip():
            assert float(line.split()[1]) < 1e-12, line
