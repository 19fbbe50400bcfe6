"""The comparator networks the flat search's first merge applies to its
register lists (gs_knn_core.hpp: 8 inputs for 7-entry lists, 16 inputs for
k = 16), read from the header's own tables: each must sort every input, which
the 0-1 principle reduces to the 2^n binary inputs (Knuth, TAOCP vol. 3,
sec. 5.3.4, Theorem Z)."""
import re
from pathlib import Path

import pytest

HDR = Path(__file__).resolve().parents[1] / "gpusteer_amd" / "csrc" / "gs_knn_core.hpp"


def _table(name: str):
    src = HDR.read_text()
    m = re.search(name + r"\[(\d+)\]\[2\]\s*=\s*\{(.*?)\};", src, re.S)
    assert m, f"table {name} not found"
    pairs = [(int(a), int(b)) for a, b in re.findall(r"\{\s*(\d+)\s*,\s*(\d+)\s*\}", m.group(2))]
    assert len(pairs) == int(m.group(1))
    return pairs


def _sorts_all_binary(net, n):
    for x in range(1 << n):
        a = [(x >> i) & 1 for i in range(n)]
        for i, j in net:
            assert i < j
            if a[j] < a[i]:
                a[i], a[j] = a[j], a[i]
        if any(a[i] > a[i + 1] for i in range(n - 1)):
            return False
    return True


@pytest.mark.parametrize("name,n,size", [("kNet", 8, 19), ("kNet16", 16, 63)])
def test_network_sorts(name, n, size):
    net = _table(name)
    assert len(net) == size
    assert all(0 <= i < n and 0 <= j < n for i, j in net)
    assert _sorts_all_binary(net, n)


def test_a_broken_network_is_caught():
    """The check itself: the 16-input network without one of its comparators
    no longer sorts."""
    net = _table("kNet16")
    assert not _sorts_all_binary(net[:30] + net[31:], 16)
