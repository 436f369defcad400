"""Host-side checks of the device helpers shared by the kernels.

The sampler reduces each 64-bit xorshift draw modulo the number of sites (or
of up / down sites) with ``mod_recip`` from ``csrc/nqs_common.h``: double
precision only, a precomputed reciprocal and one correction step, in place of
the reference's integer ``x % n`` (src/nqs/nqs_sampler.c:44-48).  The helper
is ``__host__ __device__``; here it is compiled for the host (the device
build uses the same source) and compared with ``x % n`` on random draws and
on the edge cases of the argument: multiples of n and their neighbours, the
largest 64-bit values, and moduli from 1 to 2^21 - 1.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "spin_based_neural_network_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

PROGRAM = r"""
#include "nqs_common.h"
#include <stdio.h>
int main() {
  uint64_t s = 0x9E3779B97F4A7C15ULL;
  long bad = 0, checked = 0;
  const uint32_t ns[] = {1u, 2u, 3u, 7u, 16u, 18u, 24u, 36u, 50u, 64u, 99u, 100u, 255u, 256u,
                         1000u, 65535u, 65536u, 1048575u, 2097151u};
  for (uint32_t n : ns) {
    const double rn = 1.0 / n;
    for (long t = 0; t < 400000; t++) {
      uint64_t x = nqs::xs_next(s);
      if (t < 2000) x = (uint64_t)(t / 2) * n + ((t & 1) ? 0 : n - 1);   // k n and k n - 1
      else if (t < 4000) x = ~0ULL - (uint64_t)(t - 2000);                // top of the range
      else if (t < 6000) x = ((uint64_t)(t - 4000) << 32) * n;            // high-word multiples
      checked++;
      if ((uint64_t)nqs::mod_recip(x, n, rn) != x % n) bad++;
      if ((uint64_t)nqs::mod_small(x, n) != x % n) bad++;
    }
  }
  printf("%ld %ld\n", checked, bad);
  return bad != 0;
}
"""


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not found")
def test_mod_recip_is_exact(tmp_path):
    src = tmp_path / "mod_check.hip"
    src.write_text(PROGRAM)
    exe = tmp_path / "mod_check"
    r = subprocess.run([HIPCC, "-std=c++17", "-O2", "--cuda-host-only", "-I", CSRC, str(src), "-o",
                        str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    checked, bad = map(int, out.stdout.split())
    assert checked == 19 * 400000
    assert bad == 0, f"{bad} of {checked} draws reduced differently from x % n"
