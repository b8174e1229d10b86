"""Market-data formats (include/market_data_loader.h, the reference's
include/market_data_loader.h:37-237) on the reference's own series, copied
as fixtures into tests/golden/market_data/ (tests/golden/market_data/copy_fixtures.sh):

- the C header's loaders and statistics (compiled with gcc) agree exactly with
  the Python twins in svpf_amd.data;
- the raw float32 .bin agrees with both text forms to the text's 10 decimals
  plus two float32 ulps;
- the series lengths are the reference's (market_data/market_data_info.h:8-10:
  SPY_2008_N 375, SPY_2020_N 123)."""
import os
import subprocess

import numpy as np
import pytest

from svpf_amd import data

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "market_data")
SERIES = {"spy_2008_crisis": 375, "spy_2020_covid": 123}


@pytest.fixture(scope="module")
def c_loader(tmp_path_factory):
    d = tmp_path_factory.mktemp("mdl")
    prog = d / "mdl.c"
    prog.write_text(r'''
#include "market_data_loader.h"
int main(int argc, char** argv) {
    float *r = 0, *c = 0, *lr = 0, *px = 0; int n = 0, m = 0, k = 0;
    char p[1024];
    snprintf(p, sizeof p, "%s.bin", argv[1]);
    if (load_returns_binary(p, &r, &n)) return 1;
    snprintf(p, sizeof p, "%s_returns.csv", argv[1]);
    if (load_returns_csv(p, &c, &m)) return 2;
    snprintf(p, sizeof p, "%s.csv", argv[1]);
    if (load_market_data_csv(p, &lr, &px, &k)) return 3;
    printf("%d %d %d\n", n, m, k);
    for (int i = 0; i < n; ++i) printf("%a %a %a %a\n", r[i], c[i], lr[i], px[i]);
    float mean, sd, lo, hi; int c3, c4;
    compute_return_stats(r, n, &mean, &sd, &lo, &hi, &c3, &c4);
    printf("%a %a %a %a %d %d\n", mean, sd, lo, hi, c3, c4);
    free(r); free(c); free(lr); free(px);
    return load_returns_binary("/nonexistent/x.bin", &r, &n) == -1 ? 0 : 4;
}
''')
    exe = d / "mdl"
    subprocess.run(["gcc", "-O1", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe), "-lm"], check=True)
    return exe


def _run(exe, name):
    out = subprocess.run([str(exe), os.path.join(FIX, name)], check=True, capture_output=True,
                         text=True).stdout.split("\n")
    n, m, k = map(int, out[0].split())
    vals = np.array([[float.fromhex(v) for v in ln.split()] for ln in out[1:1 + n]], dtype=np.float64)
    st = out[1 + n].split()
    stats = tuple(float.fromhex(v) for v in st[:4]) + (int(st[4]), int(st[5]))
    return (n, m, k), vals.astype(np.float32), stats


@pytest.mark.parametrize("name", sorted(SERIES))
def test_c_and_python_loaders_agree(c_loader, name):
    (n, m, k), vals, stats = _run(c_loader, name)
    assert n == m == k == SERIES[name]
    b = data.load_returns_binary(os.path.join(FIX, name + ".bin"))
    c = data.load_returns_csv(os.path.join(FIX, name + "_returns.csv"))
    lr, px = data.load_market_data_csv(os.path.join(FIX, name + ".csv"))
    assert np.array_equal(vals[:, 0], b) and np.array_equal(vals[:, 1], c)
    assert np.array_equal(vals[:, 2], lr) and np.array_equal(vals[:, 3], px)
    ps = data.return_stats(b)
    assert ps[:4] == pytest.approx(stats[:4], rel=0, abs=0) and ps[4:] == stats[4:]


@pytest.mark.parametrize("name", sorted(SERIES))
def test_binary_matches_text_forms(name):
    b = data.load_returns_binary(os.path.join(FIX, name + ".bin"))
    c = data.load_returns_csv(os.path.join(FIX, name + "_returns.csv"))
    lr, px = data.load_market_data_csv(os.path.join(FIX, name + ".csv"))
    # the text keeps 10 decimals; the .bin values sit within two float32 ulps of it
    assert np.all(np.abs(b.astype(np.float64) - c) <= 2 * np.spacing(np.abs(b)) + 1e-10)
    assert np.array_equal(c, lr)
    assert np.all(px > 0)
    # log returns are the log price ratios (the first one against the day before the file starts)
    np.testing.assert_allclose(np.log(px[1:].astype(np.float64) / px[:-1]), lr[1:], atol=2e-6)


def test_stats_by_hand():
    r = np.array([0.0, 0.01, -0.01, 0.02, -0.2], np.float32)
    m, sd, lo, hi, c3, c4 = data.return_stats(r)
    assert abs(m - r.astype(np.float64).mean()) < 1e-8 and abs(sd - r.astype(np.float64).std()) < 1e-7
    assert (lo, hi) == (float(np.float32(-0.2)), float(np.float32(0.02)))
    assert (c3, c4) == (0, 0)  # one outlier in five cannot be 3 population std away (max z = 2)
    assert data._atof("  -1.5e-3xyz") == -1.5e-3 and data._atof("") == 0.0 and data._atof("abc") == 0.0
