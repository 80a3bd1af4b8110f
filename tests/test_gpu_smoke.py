"""The smoke entry point (one small radiation call on cuda:0 checked against the oracle) runs clean."""
import pytest

pytestmark = pytest.mark.gpu


def test_smoke_entry_point():
    import __graft_entry__
    __graft_entry__.smoke()
