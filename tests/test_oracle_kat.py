"""Pins the CPU oracle with the reference's own known-answer tests
(tests/test_metal.cpp, tests/test_alpha.cpp), run through the oracle."""
import pytest

import kat_cases
import oracle_bindings as ob


def _render(exported, f, seed):
    rgb, _, _ = ob.render_frame(exported, f, seed)
    return rgb


@pytest.mark.parametrize("case", kat_cases.ALL, ids=lambda c: c.__name__)
def test_oracle_reference_kat(case):
    case(_render)
