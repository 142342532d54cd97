"""The HBM calibration probes (dnsg_hbm_probe): every kind runs and reports a
bandwidth between 1 TB/s and the MI355X's 8 TB/s peak; bad arguments are refused
before any device work."""
import pytest

from helpers import dnsg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_probe_kinds(kind):
    g = dnsg.hbm_probe(1 << 30, kind, 2)
    assert 1000.0 < g < 8000.0, g


def test_probe_rejects_bad_arguments():
    with pytest.raises(dnsg.DnsgError):
        dnsg.hbm_probe(1 << 30, 4, 2)
    with pytest.raises(dnsg.DnsgError):
        dnsg.hbm_probe(1 << 22, 3, 2)  # below 3 layers of the lattice pattern
