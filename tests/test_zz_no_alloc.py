"""The fired-list query that allocates no device memory
(scaffold_get_fired_noalloc, through the library's static scratch) against the
allocating query scaffold_get_fired, through the C ABI.  Integer results: they
must be identical."""
import numpy as np
import pytest

from conftest import new_env

pytestmark = pytest.mark.gpu


def test_fired_noalloc_matches_fired(pkg):
    big = new_env(pkg, 1)          # 96.7k cells
    big.initialize(50, 25)
    big.loop_n_steps(0, 60, 0, want_output=False)
    small = new_env(pkg, 0)        # 9.9k cells
    small.initialize(50, 25)
    small.loop_n_steps(0, 60, 0, want_output=False)
    # two calls in a row, the second with a smaller input (a step of the 9.9k-cell
    # network after one of the 96.7k-cell network), each against the allocating query
    calls = [(big, 59), (small, 59), (big, 40), (small, 12)]
    sizes = []
    for env, step in calls:
        want, n_want = env.fired(step)
        got, n_got = env.fired_noalloc(step)
        assert n_got == n_want and np.array_equal(got, want), f"step {step}"
        sizes.append(n_want)
    assert sizes[0] > sizes[1] > 0
    # a cap below the count: the first `cap` ids and the full count, as fired() gives
    cap = max(1, sizes[0] // 3)
    want, n_want = big.fired(59, cap=cap)
    got, n_got = big.fired_noalloc(59, cap=cap)
    assert n_got == n_want == sizes[0] and got.size == cap and np.array_equal(got, want)
    # sizes beyond the static scratch are refused before any device work
    with pytest.raises(pkg.ScaffoldError) as e:
        big.fired_noalloc(59, cap=(1 << 22) + 1)
    assert e.value.code == pkg.ERR_CAPACITY
    with pytest.raises(pkg.ScaffoldError) as e:   # outside the queryable steps
        big.fired_noalloc(60)
    assert e.value.code == pkg.ERR_INVALID_ARG
    big.finalize()
    small.finalize()
