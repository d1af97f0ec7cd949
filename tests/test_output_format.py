"""Host-side text formatting of the output writers (no GPU).

The swarm VTU writes coordinates as numpy's tofile(sep=" ") does
(output_swarm_to_vtu.py:72-73), and the .pvd files write the time with "%s"
(write_in_pvd_files.py:28-35): both print Python's repr of each float64.
py_repr_chars (meeuuw_amd/csrc/mw_text.h) restates that; here it is compiled
with g++ and compared against repr() and numpy's own tofile output on random
and edge-case doubles.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HDR = os.path.join(HERE, "..", "meeuuw_amd", "csrc")

HARNESS = r"""
#include "mw_text.h"
#include <cstdio>
#include <cstdint>
#include <cstring>
int main() {
  uint64_t bits;
  char out[64];
  while (fread(&bits, 8, 1, stdin) == 1) {
    double v;
    memcpy(&v, &bits, 8);
    mw::py_repr_chars(v, out);
    puts(out);
  }
  return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("fmt")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    exe = d / "h"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", HDR, str(src), "-o", str(exe)], check=True)
    return str(exe)


def _run(exe, v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    r = subprocess.run([exe], input=v.tobytes(), capture_output=True, check=True)
    return r.stdout.decode().split("\n")[:-1]


def test_py_repr_matches_python_repr(harness):
    rng = np.random.Generator(np.random.PCG64(5))
    v = np.concatenate([
        rng.uniform(0, 4, 2000),                      # particle coordinates
        rng.standard_normal(500) * 10.0 ** rng.integers(-30, 30, 500),
        np.array([0.0, -0.0, 1.0, 0.1, 1e-4, 9.999e-5, 1e-5, 1e15, 9.999999999999998e15, 1e16,
                  1.2345678901234568e16, 123456789012345.67, 5e-324, 1.7976931348623157e308,
                  -2.5, 0.5, 3.0, 100.0, 1e22, 2.0 ** -1074, np.inf, -np.inf, np.nan]),
        rng.integers(0, 10 ** 6, 200).astype(np.float64) / 8.0,  # short binary fractions
    ])
    got = _run(harness, v)
    assert got == [repr(float(x)) for x in v]


def test_py_repr_matches_numpy_tofile(harness, tmp_path):
    rng = np.random.Generator(np.random.PCG64(9))
    v = rng.uniform(0, 1, (300, 3))
    v[:, 1] = 0.0
    p = tmp_path / "t.txt"
    v.tofile(str(p), sep=" ")
    assert " ".join(_run(harness, v.ravel())) == p.read_text()


SCI_HARNESS = r"""
#include "mw_text.h"
#include <cstdio>
#include <cstdint>
#include <cstring>
int main() {
  uint64_t bits;
  char out[64];
  while (fread(&bits, 8, 1, stdin) == 1) {
    double v;
    memcpy(&v, &bits, 8);
    for (int p = 1; p <= 6; p++) {
      mw::sci_chars(v, p, out);
      fputs(out, stdout);
      fputc(p < 6 ? ' ' : '\n', stdout);
    }
  }
  return 0;
}
"""


def test_sci_matches_printf(tmp_path):
    """sci_chars(v, p) is printf's %.<p>e, the format of every VTU array and
    time series value (output_solution_to_vtu.py: %.5e, %.4e, %.3e, %.1e)."""
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "s.cpp"
    src.write_text(SCI_HARNESS)
    exe = tmp_path / "s"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", HDR, str(src), "-o", str(exe)], check=True)
    rng = np.random.Generator(np.random.PCG64(11))
    v = np.concatenate([rng.uniform(-2, 2, 3000),
                        rng.standard_normal(1000) * 10.0 ** rng.integers(-300, 300, 1000),
                        np.array([0.0, -0.0, 0.5, 1.0, 9.99995, 9.999949999, 0.125, 2.5e-5,
                                  1e100, 5e-324, 1.7976931348623157e308]),
                        np.round(rng.uniform(0, 10, 500), 5)])  # ties at the 4th/5th digit
    got = _run(str(exe), v)
    want = [" ".join(("%." + str(p) + "e") % x for p in range(1, 7)) for x in v]
    assert got == want
