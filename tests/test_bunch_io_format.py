"""The text format of dumpBunch: the reference writes with C++ ostreams in the default
float format at precisions 6 (bunch attributes), 14 (synchronous particle) and 8
(coordinates) (Bunch.cc:894-965, SyncPart.cc:465-531).  The Python writer uses printf %g at
the same precisions; this test compiles a C++ program that formats the same values through
std::ostream and checks the two agree character for character."""
import os
import shutil
import subprocess

import numpy as np
import pytest

SRC = r'''
#include <iomanip>
#include <iostream>
#include <cstdio>
int main() {
  double v; int p;
  while (std::scanf("%d %la", &p, &v) == 2) std::cout << std::setprecision(p) << v << "\n";
  return 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_python_g_format_matches_cpp_ostream(tmp_path):
    src = tmp_path / "fmt.cc"
    exe = tmp_path / "fmt"
    src.write_text(SRC)
    subprocess.run(["g++", "-O1", "-o", str(exe), str(src)], check=True)
    r = np.random.default_rng(5)
    vals = list(r.normal(size=200) * 10.0 ** r.integers(-20, 20, 200))
    vals += [0.0, -0.0, 1.0, -1.0, 0.938272013, 1.534698e-18, 1e8, 1e-5, 123456789.0,
             0.1, 1.0 / 3.0, 2.5e-300, 1.7976931348623157e308, 5e-324, 1e16, 1e15, 999999.5]
    cases = [(p, v) for v in vals for p in (6, 8, 14)]
    inp = "".join(f"{p} {float(v).hex()}\n" for p, v in cases)
    out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True)
    cpp = out.stdout.split("\n")[:-1]
    py = [("%." + str(p) + "g") % v for p, v in cases]
    assert cpp == py
