"""The library's host-side SCC pieces through the C ABI on the CPU (no GPU needed): the Fermi
filling (dftbhip_efilling, Efilling of etemp.F90:59-194) and the Broyden mixer
(dftbhip_mixer_*, TBroydenMixerReal of broydenmixer.F90:112-342), against the reference's own
regression sequence and the oracle."""
import json
import math
import os

import numpy as np
import pytest

from dftbplus_amd import api
from oracle import dftb_ref as R

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
EPS = np.finfo(float).eps


def test_library_broyden_reproduces_reference_sequence():
    """test/src/dftbp/unit/mixer/broyden.F90:24-58, to the reference's 100 eps."""
    fx = json.load(open(os.path.join(GOLDEN, "broyden_h2o.json")))
    ch = np.array(fx["charges"])
    df = np.array(fx["differences"])
    mix = api.BroydenMixer(3, fx["mix_param"], fx["omega0"], fx["min_weight"], fx["max_weight"],
                           fx["weight_fac"], max_scc_iter=100)
    q = ch[0].copy()
    for i in range(4):
        q = mix.mix(q, df[i])
        assert np.max(np.abs(q - ch[i + 1])) < 100 * EPS, i
    mix.close()


def test_library_broyden_equals_oracle_on_a_long_sequence():
    """Twenty steps of a damped linear fixed-point problem q -> M q + b: the library's mixer and
    the oracle's agree to rounding at every step (history, weights and the LU solve)."""
    rng = np.random.default_rng(11)
    n = 37
    M = rng.standard_normal((n, n)) * 0.06
    b = rng.standard_normal(n)
    q_lib = np.zeros(n)
    q_ref = np.zeros(n)
    lib_mix = api.BroydenMixer(n, 0.2, 0.01, 1.0, 1e5, 1e-2, max_scc_iter=30)
    ref_mix = R.BroydenMixer(n, 0.2, 0.01, 1.0, 1e5, 1e-2, 31)
    for it in range(20):
        d_lib = M @ q_lib + b - q_lib
        d_ref = M @ q_ref + b - q_ref
        q_lib = lib_mix.mix(q_lib, d_lib)
        q_ref = ref_mix.mix(q_ref, d_ref)
        assert np.max(np.abs(q_lib - q_ref)) < 1e-12 * max(1.0, np.max(np.abs(q_ref))), it
    # the mixing converges towards the fixed point
    r0 = np.max(np.abs(b))
    assert np.max(np.abs(M @ q_lib + b - q_lib)) < 1e-4 * r0
    lib_mix.close()


def test_library_broyden_step_limit_is_an_error():
    """More mix calls than maxSccIter after a reset raise, as the reference does."""
    mix = api.BroydenMixer(2, max_scc_iter=3)
    q = np.array([1.0, 2.0])
    for _ in range(3):
        q = mix.mix(q, np.array([0.1, -0.1]))
    with pytest.raises(api.DftbHipError, match="Maximal nr. of steps exceeded"):
        mix.mix(q, np.array([0.1, -0.1]))
    mix.close()


def _oracle_one_spin(eig, nel_spin, kT, kw=None):
    """Oracle fillings for one spin channel (its spin-unpolarised routines carry a factor 2)."""
    if kw is None:
        f, ef, eb, ts, e0 = R.fermi_filling(eig, 2 * nel_spin, kT)
    else:
        f, ef, eb, ts, e0 = R.fermi_filling_k(eig, 2 * nel_spin, kT, kw)
    return 0.5 * f, ef, 0.5 * eb, 0.5 * ts, 0.5 * e0


@pytest.mark.parametrize("kT,nel", [(1e-8, 20.0), (0.01, 31.0), (0.05, 17.3), (0.01, 0.0),
                                    (0.01, 50.0), (1e-3, 24.0)])
def test_library_efilling_equals_oracle_gamma(kT, nel):
    rng = np.random.default_rng(2)
    eig = np.sort(rng.uniform(-1, 1, 50))
    f, ef, eb, ts, e0 = api.efilling(eig, np.array([1.0]), nel, kT)
    rf, ref_ef, reb, rts, re0 = _oracle_one_spin(eig, nel, kT)
    np.testing.assert_allclose(f[0], rf, rtol=1e-13, atol=1e-15)
    assert ef == pytest.approx(ref_ef, rel=1e-14, abs=1e-14)
    for a, b in ((eb, reb), (ts, rts), (e0, re0)):
        assert a == pytest.approx(b, rel=1e-13, abs=1e-14)
    if nel == 0.0:
        assert np.all(f == 0.0) and eb == 0.0
        assert ef == eig[0] - 1000.0 * (kT + float(np.finfo(np.float32).eps))


def test_library_efilling_rescale_branch():
    """The final rescale to the exact count (etemp.F90:186-189), as pinned for the oracle."""
    eig = np.array([-0.8, -0.5, 0.3, 0.3, 0.3, 0.9, 1.4])
    f, ef, eb, ts, e0 = api.efilling(eig, np.array([1.0]), 2 + 1.37, 1e-8)
    assert abs(f.sum() - (2 + 1.37)) < 1e-13
    assert f[0, 2] == f[0, 3] == f[0, 4]
    x = 1.37 / 3
    assert abs(ef - (0.3 - 1e-8 * math.log(1.0 / x - 1.0))) < 1e-14


@pytest.mark.parametrize("nel", [0.0, 31.0, 30.4])
def test_library_efilling_equals_oracle_kpoints(nel):
    rng = np.random.default_rng(5)
    nk, nlev = 4, 40
    eig = np.sort(rng.uniform(-1, 1, (nk, nlev)), axis=1)
    kw = np.array([0.125, 0.375, 0.25, 0.25])
    f, ef, eb, ts, e0 = api.efilling(eig, kw, nel, 0.01)
    rf, ref_ef, reb, rts, re0 = _oracle_one_spin(eig, nel, 0.01, kw)
    np.testing.assert_allclose(f, rf, rtol=1e-13, atol=1e-15)
    assert ef == pytest.approx(ref_ef, rel=1e-14, abs=1e-14)
    assert eb == pytest.approx(reb, rel=1e-13, abs=1e-14)
    assert ts == pytest.approx(rts, rel=1e-13, abs=1e-14)


def test_library_efilling_argument_errors():
    with pytest.raises(api.DftbHipError, match="ascending"):
        api.efilling(np.array([0.2, 0.1]), np.array([1.0]), 1.0, 0.01)
    with pytest.raises(api.DftbHipError, match="too many electrons"):
        api.efilling(np.array([0.1, 0.2]), np.array([1.0]), 5.0, 0.01)
