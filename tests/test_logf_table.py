"""The exact-logf correction table (film_grain_amd/csrc/fg_logf.h): the
generator checks, over every float in [1e-6, 1], that the table corrects the
float of fg_log_double to this machine's C library logf (the reference's
f32::ln), and that no correction exceeds one unit in the last place.  Here it
runs again on the CPU; the device side is checked over the same domain by
tests/test_gpu_frame.py."""
import re

from film_grain_amd import build


def test_generator_verifies_the_whole_domain():
    header = build.logf_table()
    text = header.read_text()
    n = int(re.search(r"kLogfTableN = (\d+);", text).group(1))
    entries = re.findall(r"0x[0-9a-f]{8}u", text)
    assert len(entries) == n
    keys = [int(e[:-1], 16) & 0x7FFFFFFF for e in entries]
    assert keys == sorted(keys) and len(set(keys)) == n
    # glibc's logf is correctly rounded except on about 0.1% of the domain
    assert 0 < n < 1_000_000
