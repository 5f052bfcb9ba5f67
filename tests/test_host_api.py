"""Host-side argument checks of the solver entry points, before any device call
(converge!'s keywords, src/methods/converge.jl:51-60), and the progress hook's C
type.  No GPU."""
import numpy as np
import pytest

from cdft_amd import _lib, configs, converge_


@pytest.fixture(scope="module")
def wall():
    return configs.methane_wall()


def test_converge_keyword_checks(wall):
    s, r = wall
    with pytest.raises(ValueError):
        converge_(s, r.copy(), log_interval=-1)
    with pytest.raises(ValueError):
        converge_(s, r.copy(), save_interval=-3, save_callback=lambda it, x: None)
    with pytest.raises(ValueError):                      # DFTProblem asserts a callback
        converge_(s, r.copy(), save_interval=2)
    with pytest.raises(TypeError):
        converge_(s, r.astype(np.float32), save_interval=2, save_callback=lambda it, x: None)


def test_progress_hook_type_matches_the_header():
    """int (*)(void *user, int iter, const double *rho, long n); the options carry it
    after the history fields (the layout itself is checked against gcc in test_abi)."""
    f = _lib.ProgressFn(lambda u, it, rho, n: 0)
    assert f(None, 3, None, 10) == 0
    names = [n for n, _ in _lib.SolverOpts._fields_]
    assert names[-3:] == ["progress", "progress_user", "progress_every"]
    assert _lib.CDFT_ERR_STOPPED == -5
    o = _lib.SolverOpts(10, 0.01, 1e-8, 0.1, 5, 0, None)
    assert not o.progress and o.progress_every == 0


def test_new_entry_points_are_declared_in_the_header():
    """cdft_adjoint_params and the progress hook are part of the C ABI header as well
    as of the Python binding's export list (test_abi checks the export list against
    the built library)."""
    import os
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "cdft_hip.h")).read()
    assert "int cdft_adjoint_params(" in hdr and "cdft_adjoint_params" in _lib.EXPORTS
    assert "typedef int (*cdft_progress_fn)(void *user, int iter, const double *rho, long n);" in hdr
    assert "#define CDFT_ERR_STOPPED (-5)" in hdr
