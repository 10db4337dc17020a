"""The tests' environment-switch helper (tests/_cases.py, used by the GPU
tests that force kernel forms): values set for the block, None unsets, and
the previous values come back afterwards, also when the block raises."""
import os

import pytest

from _cases import env


def test_env_sets_unsets_and_restores():
    os.environ["SCFT_TEST_A"] = "old"
    os.environ.pop("SCFT_TEST_B", None)
    try:
        with pytest.raises(RuntimeError):
            with env(SCFT_TEST_A=None, SCFT_TEST_B="1"):
                assert "SCFT_TEST_A" not in os.environ
                assert os.environ["SCFT_TEST_B"] == "1"
                raise RuntimeError("leave the block")
        assert os.environ["SCFT_TEST_A"] == "old"
        assert "SCFT_TEST_B" not in os.environ
    finally:
        os.environ.pop("SCFT_TEST_A", None)
        os.environ.pop("SCFT_TEST_B", None)
