"""The spectral path's autotuner: every line-FFT variant it may pick (lines per
tile, the all-pairs z pass) gives the same functional and the same solve bit for
bit, and a tuned choice is valid, measured and shared by systems of one shape.
Needs a GPU."""
import sys

import numpy as np
import pytest

from cdft_amd import DFTOptions, configs, converge_
from cdft_amd._lib import CDFTError
from helpers import smooth_profile

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def variants(nz_fused):
    for tb in (4, 8, 16):
        for fz in ((1, 0) if nz_fused else (0,)):
            yield tb, fz


@pytest.mark.parametrize("shape", ["3d", "2d"])
def test_variants_are_bitwise_identical(shape):
    mixed = DFTOptions(precision="mixed")
    if shape == "3d":                              # z lines of 128: the all-pairs z pass is a variant
        from cdft_amd import DFTSystem, PCSAFT, Uniform3DCart
        m = PCSAFT(["methane", "ethane"])
        T, p, x = 200.0, 1e7, np.array([0.7, 0.3])
        v = m.volume(p, T, x, "liquid")
        L = m.length_scale()
        s = DFTSystem(m, Uniform3DCart((p, T), x / v, [[0, 6 * L], [0, 6 * L], [0, 24 * L]], (32, 32, 128)),
                      None, mixed)
    else:
        s, _ = configs.matrix_2d(n=64, box=10.0, ndisks=2, seed=3, options=mixed)
    dev = s.device_system
    rho = smooth_profile(s, seed=2).reshape(-1)
    fused = shape == "3d"
    ref = None
    for tb, fz in variants(fused):
        dev.set_tuning((tb, tb), (fz, fz))
        got = (dev.functional_eval(rho)[0], dev.functional_eval_fp32(rho))
        r = rho.copy()
        converge_(s, r, tol=1e-9, maxit=4)
        got = got + (r,)
        say(f"{shape}: lines per tile {tb}, fused z {fz}: |dF/drho| {np.max(np.abs(got[0])):.6e}")
        if ref is None:
            ref = got
        else:
            for a, b in zip(got, ref):
                assert np.array_equal(a, b), (tb, fz)


def test_tune_picks_a_valid_measured_choice_shared_by_shape():
    from cdft_amd import DFTSystem, Uniform3DCart
    s, _ = configs.matrix_3d(n=64, box=8.0, nspheres=2, seed=1, options=DFTOptions(precision="mixed"))
    t = s.device_system.tune(force=True)
    say(f"64^3 mixed: {t}")
    assert t["tuned"] == 1
    for pr in range(2):
        assert t["lines_per_tile"][pr] in (4, 8, 16) and t["ms"][pr] > 0
    s2, _ = configs.matrix_3d(n=64, box=8.0, nspheres=3, seed=4, options=DFTOptions(precision="mixed"))
    t2 = s2.device_system.tune()
    assert t2 == t                                  # the shape's cached choice, not measured again
    with pytest.raises(CDFTError):
        s2.device_system.set_tuning((32, 8), (0, 0))


def test_one_dimensional_system_has_nothing_to_tune():
    s, _ = configs.methane_wall()
    t = s.device_system.tune(force=True)
    assert t["tuned"] == 0 and t["lines_per_tile"] == [8, 8]
