"""The SYMV tile autotuner through the C ABI: a well-formed table whose buckets took their fastest
option, and tridiagonalisations under the installed table that keep the spectrum."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPTIONS = (2, 4, 8, 16)


def _check_rows(rows, secs, n):
    """Table invariants: ascending distinct starts at 0, valid options, no repeated neighbours."""
    assert rows and rows[0][0] == 0
    starts = [r[0] for r in rows]
    assert starts == sorted(starts) and len(set(starts)) == len(starts)
    assert all(t in OPTIONS for _, t in rows)
    assert all(rows[k][1] != rows[k + 1][1] for k in range(len(rows) - 1))
    nb = (n + 255) // 256
    assert secs.shape == (4, nb) and np.all(secs > 0)
    # every bucket's option is within 2 % of its fastest; a switch is to a bucket's fastest
    for b in range(nb):
        tiles = [t for m0, t in rows if m0 <= 256 * b][-1]
        assert secs[OPTIONS.index(tiles), b] <= np.min(secs[:, b]) / 0.98 * (1 + 1e-12)
    for m0, t in rows:
        assert t == OPTIONS[int(np.argmin(secs[:, m0 // 256]))]


@pytest.mark.parametrize("precision", [0, 1])
def test_autotune_table(precision):
    from dftbplus_amd import api
    n = 1536
    try:
        rows, secs = api.autotune(n, precision)
        print(f"precision {precision}: rows {rows}")
        print("seconds per option (rows) and bucket (columns):\n", np.array2string(secs, precision=5))
        _check_rows(rows, secs, n)
    finally:
        api.set_tuning(1, 0)


def test_tridiagonalisation_under_tuned_table():
    from dftbplus_amd import api
    rng = np.random.default_rng(5)
    n = 1100
    a = rng.standard_normal((n, n))
    a = 0.5 * (a + a.T)
    ref = np.linalg.eigvalsh(a)
    try:
        rows, _ = api.autotune(1024, 0)
        _, d, e, _ = api.dsytrd(a)
        t = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        w = np.linalg.eigvalsh(t)
        assert np.max(np.abs(w - ref)) < 1e-11 * np.max(np.abs(ref))
    finally:
        api.set_tuning(1, 0)
    with pytest.raises(api.DftbHipError, match="n must be"):
        api.autotune(100, 0)
    with pytest.raises(api.DftbHipError, match="precision"):
        api.autotune(1024, 2)
