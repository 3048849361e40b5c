"""Pin of the oracle's hkl family ordering against libstdc++'s own
std::stable_sort.  The reference's comparator (d within 1e-6, then
|F|^2*mult within 1e-6, multiplicity, hkl) is not a strict weak ordering, so
the order of near-tied families depends on the sort algorithm itself; the
oracle restates that algorithm and this test compiles a few lines of C++
with the system g++ (the toolchain the reference is built with) to compare.
CPU only."""
import pathlib
import shutil
import subprocess

import numpy as np
import pytest

CPP = r"""
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>
struct F { double d, f; int m, h, k, l, idx; };
static bool hkl_less(const F& a, const F& o) {
  return (a.h != o.h ? o.h < a.h : (a.k != o.k ? o.k < a.k : o.l < a.l));
}
int main() {
  int n;
  if (scanf("%d", &n) != 1) return 1;
  std::vector<F> v(n);
  for (int i = 0; i < n; ++i) {
    scanf("%lf %lf %d %d %d %d", &v[i].d, &v[i].f, &v[i].m, &v[i].h, &v[i].k, &v[i].l);
    v[i].idx = i;
  }
  std::stable_sort(v.begin(), v.end(), [](const F& rh, const F& lh) {
    if (std::fabs(lh.d - rh.d) > 1e-6) return lh.d < rh.d;
    const double lfm = lh.f * lh.m, rfm = rh.f * rh.m;
    if (std::fabs(lfm - rfm) > 1e-6) return lfm < rfm;
    if (lh.m != rh.m) return lh.m < rh.m;
    return hkl_less(lh, rh);
  });
  for (auto& x : v) printf("%d\n", x.idx);
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("n", [1, 2, 7, 8, 15, 29, 1000, 20001])
def test_family_order_matches_libstdcxx(oracle, tmp_path, n):
    exe = tmp_path / "ss"
    src = tmp_path / "ss.cc"
    src.write_text(CPP)
    subprocess.run(["g++", "-O2", "-o", str(exe), str(src)], check=True)
    rng = np.random.default_rng(n)
    # clustered d values: many pairs closer than the 1e-6 tolerance, chains
    # of near-ties that make the comparator non-transitive
    base = rng.uniform(0.2, 10.0, size=(n + 4) // 5)
    d = np.repeat(base, 5)[:n] + rng.integers(0, 4, size=n) * 6e-7
    rng.shuffle(d)
    f = rng.choice([0.5, 1.0, 1.0 + 4e-7, 2.0], size=n)
    m = rng.choice([2, 4, 8], size=n).astype(np.int32)
    hkl = rng.integers(-3, 4, size=(n, 3)).astype(np.int32)
    lines = [str(n)] + [f"{d[i]:.17g} {f[i]:.17g} {m[i]} {hkl[i,0]} {hkl[i,1]} {hkl[i,2]}" for i in range(n)]
    out = subprocess.run([str(exe)], input="\n".join(lines), capture_output=True, text=True, check=True).stdout
    ref = np.array([int(x) for x in out.split()], dtype=np.int32)
    got = oracle.debug_sort_fams(d, f, m, hkl)
    assert np.array_equal(got, ref)
