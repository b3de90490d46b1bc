"""The exponential the device uses in place of glibc's exp (fg_exp.h), on the CPU.

* fg_exp_tables.h is what gen_exp_tables.py writes (2^(j/64) as
  double-doubles from 80-digit decimals, ln2/64 split with a 32-bit high
  part), and its values are the nearest doubles.
* The build's glibc table (fg_exp_neg_table.h) was checked exhaustively by
  its generator; here its size is bounded (glibc's exp misrounds about one
  input in 10^4) and its entries are ascending, as the lookup's binary search
  needs.
"""
import re
import subprocess
import sys
from decimal import Decimal, getcontext
from fractions import Fraction
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "film_grain_amd" / "csrc"


def _values(txt, name):
    body = txt.split(f"#define {name}")[1].split("\n#")[0]
    return [float(v) for v in re.findall(r"-?[0-9][0-9.e+-]*", body)]


def test_constants_regenerate_identically(tmp_path):
    import shutil
    shutil.copy(CSRC / "gen_exp_tables.py", tmp_path / "gen_exp_tables.py")
    subprocess.run([sys.executable, str(tmp_path / "gen_exp_tables.py")], check=True)
    assert (tmp_path / "fg_exp_tables.h").read_text() == (CSRC / "fg_exp_tables.h").read_text()


def test_constants_are_nearest_doubles():
    getcontext().prec = 60
    txt = (CSRC / "fg_exp_tables.h").read_text()
    th, tl = _values(txt, "FG_EXP_T_HI_VALUES"), _values(txt, "FG_EXP_T_LO_VALUES")
    assert len(th) == len(tl) == 64
    for j in range(64):
        exact = Decimal(2) ** (Decimal(j) / 64)
        # hi is within half a unit of the exact value, hi + lo within 2^-100
        assert abs(Decimal(th[j]) - exact) <= Decimal(2) ** -53 * exact
        assert abs(Decimal(th[j]) + Decimal(tl[j]) - exact) <= Decimal(2) ** -100
    l_hi = float(re.search(r"FG_EXP_L_HI (\S+)", txt).group(1))
    l_lo = float(re.search(r"FG_EXP_L_LO (\S+)", txt).group(1))
    m = Fraction(l_hi)
    while m.denominator > 1:
        m *= 2
    assert int(m).bit_length() - (int(m) & -int(m)).bit_length() + 1 <= 32, "L_HI has more than 32 bits"
    assert abs(Decimal(l_hi) + Decimal(l_lo) - Decimal(2).ln() / 64) <= Decimal(2) ** -90


def test_glibc_table_is_small_and_sorted():
    hdr = ROOT / "film_grain_amd" / "_build" / "fg_exp_neg_table.h"
    if not hdr.exists():
        from film_grain_amd import build
        build.BUILD.mkdir(exist_ok=True)
        build.exp_table()
    txt = hdr.read_text()
    n = int(re.search(r"FG_EXP_NEG_TABLE_N (\d+)", txt).group(1))
    vals = [int(v, 16) & 0x7FFFFFFF for v in re.findall(r"0x([0-9a-f]{8})u", txt)]
    assert len(vals) == n
    assert 0 < n < 1.1e-3 * 0x41400000
    assert all(a < b for a, b in zip(vals, vals[1:]))
    assert vals[-1] < 0x41400000
