"""The plausibility limits the device checks after every step (include/his_amd.h) are the ones the
host's check_plausible uses (hydrointel/solver/coupling.py:169-192: 10 m of 2-D depth, 10 m of
channel level above the bank), and every device error code has a message."""
import re
from pathlib import Path

from his_amd.solver import coupling

HEADER = Path(__file__).resolve().parents[1] / "include" / "his_amd.h"


def _define(name):
    m = re.search(rf"#define\s+{name}\s+([0-9.eE+-]+)", HEADER.read_text())
    assert m, name
    return float(m.group(1))


def test_device_limits_equal_host_limits():
    assert _define("HIS_PLAUSIBLE_DEPTH_2D_M") == coupling.MAX_PLAUSIBLE_DEPTH_2D_M == 10.0
    assert _define("HIS_PLAUSIBLE_RISE_1D_M") == coupling.MAX_PLAUSIBLE_RISE_1D_M == 10.0


def test_every_device_error_code_has_a_message():
    text = HEADER.read_text()
    block = text[text.index("int32_t *error;"):text.index("int32_t n_sub_limit;")]
    codes = {int(c) for c in re.findall(r"(\d) = ", block)}
    assert codes == set(coupling.ERROR_CODES) == {1, 2, 3, 4, 5}


def test_network_part_limit_matches_the_library():
    from his_amd.solver import swe1d
    assert _define("HIS_MAX_NET_PARTS") == swe1d.MAX_NET_PARTS == 63
    assert swe1d.MAX_PARTS <= swe1d.MAX_NET_PARTS
