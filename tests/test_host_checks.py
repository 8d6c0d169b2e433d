"""Host-side argument checks that need no GPU."""
import numpy as np
import pytest


def test_observed_traces_shape_checks():
    from viscoelasticinversion_amd.stepper import observed_traces
    d = np.arange(3 * 50, dtype=np.float64).reshape(3, 50)
    nt, out = observed_traces(d, 3, 40, np.float32)
    assert nt == 40 and out.shape == (3, 40) and out.dtype == np.float32
    assert out.flags.c_contiguous and np.array_equal(out, d[:, :40].astype(np.float32))
    nt, out = observed_traces(d, 3, None, np.float64)
    assert nt == 50 and out.shape == (3, 50)
    with pytest.raises(ValueError, match="fewer than nt"):
        observed_traces(d, 3, 60, np.float32)          # too short
    with pytest.raises(ValueError, match="receivers"):
        observed_traces(d.T, 3, 3, np.float32)         # transposed (nt, nrec)
    with pytest.raises(ValueError, match="2-D"):
        observed_traces(d.ravel(), 3, 40, np.float32)  # flat
    with pytest.raises(ValueError, match="nt must be"):
        observed_traces(d, 3, 0, np.float32)


def test_step_parameters_type_is_checked():
    from viscoelasticinversion_amd.api import apply_parameters
    with pytest.raises(TypeError):
        apply_parameters(object(), {"lam": 1.0})
