"""The Python wrappers refuse state arrays whose shapes do not match the
cells they describe, before any C call (the C side reads exactly n cells
from each array).  No GPU is needed: the checks run first."""
import numpy as np
import pytest

from proteusgpu_amd import api


def test_state_arrays_match_cells():
    rho, vel, E = api._state_arrays(np.ones(4), np.zeros((4, 3)), np.ones(4), 4, 3)
    assert rho.dtype == np.float64 and vel.shape == (4, 3) and E.flags["C_CONTIGUOUS"]
    for bad in [(np.ones(3), np.zeros((4, 3)), np.ones(4)), (np.ones(4), np.zeros((4, 2)), np.ones(4)),
                (np.ones(4), np.zeros((4, 3)), np.ones(5)), (np.ones(4), np.zeros(12), np.ones(4))]:
        with pytest.raises(ValueError):
            api._state_arrays(*bad, 4, 3)


def test_hll_flux_shapes():
    L = np.ones((5, 5))
    with pytest.raises(ValueError):
        api.hll_flux(None, 1.4, L, np.ones((4, 5)), np.ones((5, 3)))
    with pytest.raises(ValueError):
        api.hll_flux(None, 1.4, L, L, np.ones((5, 2)))
    with pytest.raises(ValueError):
        api.hll_flux(None, 1.4, L, L, np.ones((5, 3)), wall=np.zeros(4))
    with pytest.raises(ValueError):
        api.hll_flux(None, 1.4, np.ones((5, 6)), np.ones((5, 6)), np.ones((5, 3)))
