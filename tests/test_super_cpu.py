"""Segment arithmetic of the super operator (operators.h:562-576): seg_num blocks of floor(M / seg_num)
visibilities, the last taking the remainder."""
import pytest

from purify_amd.operators import _segments


def test_segments_split_like_the_reference():
    assert _segments(10, 3) == [(0, 3), (3, 3), (6, 4)]
    assert _segments(9, 3) == [(0, 3), (3, 3), (6, 3)]
    assert _segments(2, 3) == [(0, 0), (0, 0), (0, 2)]
    assert sum(b for _, b in _segments(13107, 7)) == 13107


def test_zero_segments_raise():
    with pytest.raises(RuntimeError, match="Number of segments = 0"):
        _segments(10, 0)
