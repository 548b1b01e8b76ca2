"""Host-side logic checked without a GPU."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_division_free_opening_test_is_exact(tmp_path):
    # the traversal decides every node with  r >= rmin0 * 2^layer, or (quadrupoles, fast division)
    # r2 >= c0 * 4^layer  (astro_amd/csrc/accept.hpp); both must agree with the reference's
    # side / r < theta  evaluated as an IEEE division of r = RN(sqrt(r2))
    exe = tmp_path / "accept_check"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                           os.path.join(ROOT, "tests", "cpp", "accept_check.cpp"), "-o", str(exe)])
    bad, near, total, bad2, total2 = map(int, subprocess.check_output([str(exe)]).split())
    assert total == 2000000 and near == 1000000  # 10 thetas x 200,000
    assert bad == 0
    # the squared form  r2 >= accept_r2min(rmin0) 4^l
    assert total2 > 1900000
    assert bad2 == 0
