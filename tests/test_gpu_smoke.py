"""The package's smoke entry point (one FP64 step on cuda:0 against the
oracle, and the sort's permutation against the reference's stable counting
sort) run as a test, so that it keeps working."""
import pytest

pytestmark = pytest.mark.gpu


def test_smoke(gpu):
    import __graft_entry__
    __graft_entry__.smoke()
