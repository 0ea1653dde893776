"""Chemistry runtime mirrors: argument handling and index conventions (CPU)."""
from __future__ import annotations

import pytest

from tyxonq_amd import runtimes as R


def test_spin_orbital_terms_reverse_the_index():
    of = {((0, "X"), (3, "Z")): 0.5, (): -1.25, ((2, "Y"),): 0.75}
    got = R.spin_orbital_terms(of, 4)
    assert (0.5, [("X", 3), ("Z", 0)]) in got
    assert (-1.25, []) in got
    assert (0.75, [("Y", 1)]) in got

    class QOp:
        terms = {((1, "z"),): 2.0}

    assert R.spin_orbital_terms(QOp(), 3) == [(2.0, [("Z", 1)])]
    with pytest.raises(ValueError):
        R.spin_orbital_terms({((0, "X"),): 1j}, 2)


def test_unsupported_engines_and_modes_raise():
    with pytest.raises(NotImplementedError):
        R._engine_name("civector")
    assert R._engine_name(None) == "statevector"
