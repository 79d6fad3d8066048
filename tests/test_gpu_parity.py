"""GPU parity: the HIP engine against the CPU oracle on the same seeded inputs.

Integer and byte outputs (spike rasters, refractory counters, counts, events,
fired lists) must match exactly; V and g are compared bit for bit as well,
because the engine evaluates the reference's FP32 expressions in the same order
without contraction (tolerance 0 ulp)."""
import numpy as np
import pytest

from conftest import new_env

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def assert_same_state(env, ref, tag=""):
    V, g, r = env.state()
    V0, g0, r0 = ref.state()
    bad = np.nonzero(bits(V) != bits(V0))[0]
    assert bad.size == 0, f"{tag}: V differs at {bad[:10]} gpu={V[bad[:5]]} oracle={V0[bad[:5]]}"
    bad = np.nonzero(bits(g) != bits(g0))[0]
    assert bad.size == 0, f"{tag}: g differs at {bad[:10]} gpu={g[bad[:5]]} oracle={g0[bad[:5]]}"
    assert np.array_equal(r, r0), f"{tag}: refractory counters differ"


def outdegree(env):
    info = env.info()
    deg = np.zeros(info.total_neurons, np.uint64)
    for i in range(info.n_projections):
        p = env.projection(i)
        if p.pre < 0:
            continue
        _, col, _ = env.projection_csr(i)
        base = env.population(p.pre).base_id
        np.add.at(deg, base + col.astype(np.int64), 1)
    return deg


def run_and_compare(env, ref, n0, steps, chunk=25, check_counts=True):
    info = env.info()
    bases = [env.population(i).base_id for i in range(info.n_populations)]
    nums = [env.population(i).num for i in range(info.n_populations)]
    deg = outdegree(env) if check_counts else None
    for s0 in range(n0, n0 + steps, chunk):
        env.loop_n_steps(s0, chunk, 0, want_output=False)
        ref.run(s0, chunk)
        for s in range(s0, s0 + chunk):
            a, b = env.spike_row(s), ref.spike_row(s)
            if not np.array_equal(a, b):
                d = np.nonzero(a != b)[0]
                raise AssertionError(f"raster differs at step {s}: ids {d[:10]} gpu={a[d[:10]]} oracle={b[d[:10]]}")
            if check_counts:
                c, ev = env.counts(s)
                want = [int(b[bases[p]:bases[p] + nums[p]].sum()) for p in range(len(bases))]
                assert list(c) == want, f"counts differ at step {s}: {list(c)} vs {want}"
                assert ev == int(deg[b.astype(bool)].sum()), f"events differ at step {s}"


def test_smoke():
    import __graft_entry__ as g
    g.smoke()


@pytest.mark.parametrize("graph", [0, 1])
def test_config0_bit_exact(pkg, orc, graph):
    env = new_env(pkg, 0)
    env.set_option("graph", graph)
    env.initialize(50, 25)
    ref = orc.OracleSim.from_env(env)
    ref.initialize()
    assert_same_state(env, ref, "init")
    run_and_compare(env, ref, 0, 400)
    assert_same_state(env, ref, "after 400 steps")
    env.finalize()


def test_config1_bit_exact(pkg, orc):
    env = new_env(pkg, 1)
    env.initialize(50, 25)
    ref = orc.OracleSim.from_env(env)
    ref.initialize()
    run_and_compare(env, ref, 0, 150, check_counts=True)
    assert_same_state(env, ref, "config1 150 steps")
    env.finalize()
