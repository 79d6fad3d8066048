"""Host side of the C ABI, without a GPU: symbol exports, registration order,
input validation, the .dat loader, options and the synthetic scaffold generator.

None of these tests launches a kernel; the engine is only created by
scaffold_initialize_network_env, which these tests either avoid or expect to
fail with SCAFFOLD_ERR_NO_DEVICE on a machine without an MI355X."""
import ctypes
import os

import numpy as np
import pytest

from conftest import new_env

GRANULE = (3.0, -74.0, 1.5, 0.0, -84.0, -42.0, 1.5, (1., 0., 1.0, 0.5, 1., -85., 1.0, 10.0))
HAS_GPU_NODE = os.path.exists("/dev/kfd")


def test_library_exports_every_header_symbol(pkg):
    lib = pkg.load_library()
    names = pkg.header_symbols()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    # and nothing beyond the header is exported under the library prefix
    import shutil
    import subprocess
    if shutil.which("nm"):
        out = subprocess.run(["nm", "-D", "--defined-only", pkg.LIB_PATH], capture_output=True, text=True).stdout
        exported = {ln.split()[-1] for ln in out.splitlines() if " T scaffold_" in ln}
        assert exported == set(names)


def test_version_and_error_text(pkg):
    lib = pkg.load_library()
    lib.scaffold_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.scaffold_version()
    with pytest.raises(pkg.ScaffoldError) as e:
        pkg.NetworkEnvironment(2, 7, 16, 2, 0)
    assert e.value.code == pkg.ERR_UNSUPPORTED and "one GPU" in str(e.value)


@pytest.mark.parametrize("args,code", [((1, 0, 1, 1, 0), "ERR_INVALID_ARG"), ((1, 7, -1, 1, 0), "ERR_INVALID_ARG"),
                                       ((1, 7, 16, 2, -1), "ERR_INVALID_ARG"), ((1, 33, 1, 1, 0), "ERR_INVALID_ARG")])
def test_create_validation(pkg, args, code):
    with pytest.raises(pkg.ScaffoldError) as e:
        pkg.NetworkEnvironment(*args)
    assert e.value.code == getattr(pkg, code)


def test_reference_stimulus_functions(pkg):
    bg = pkg.stim_func_reference(0, 5)
    assert bg.poisson_hz == 1.0 and bg.poisson_on == 0.0 and bg.poisson_off == float("inf")
    assert bg.periodic_hz == 0.0 and bg.seed == 5
    tone = pkg.stim_func_reference(1, 6)
    assert tone.poisson_hz == 1.0 and tone.poisson_on == -float("inf")
    assert tone.periodic_hz == 140.0 and np.float32(tone.periodic_start) == np.float32(292.8)
    assert tone.periodic_on == 300.0 and tone.periodic_off == 351.0 and tone.periodic_off_inclusive == 0
    assert bg.sine_max_hz == 0.0 and tone.sine_max_hz == 0.0
    puff = pkg.stim_func_reference(2, 7)   # puff_stim, UserInputFunctions.cu:33-39
    assert puff.poisson_hz == 0.0 and puff.burst_hz == 0.0 and puff.periodic_hz == 0.0
    assert (puff.sine_max_hz, puff.sine_cycle_s, puff.sine_shift) == (3.0, 2.0, 0.0)
    assert puff.sine_on == 0.0 and puff.sine_off == float("inf") and puff.seed == 7
    with pytest.raises(pkg.ScaffoldError):
        pkg.stim_func_reference(3, 0)


@pytest.mark.parametrize("cycle,shift,ok", [(2.0, 0.0, True), (0.0, 0.0, False), (float("inf"), 0.0, False),
                                            (2.0, float("nan"), False)])
def test_sinusoidal_stimulus_validation(pkg, cycle, shift, ok):
    env = pkg.NetworkEnvironment(1, 1, 0, 1, 0)
    ch = (1., 0., 1.0, 0.5, 1., -85., 1.0, 10.0)
    g = env.set_neuron_params(0, "glom", 0, 4, 3.0, -74.0, 1.5, 0.0, -84.0, -42.0, 1.5, pkg.INPUT, ch)
    f = pkg.stim_func_reference(2, 1)
    f.sine_cycle_s, f.sine_shift = cycle, shift
    if ok:
        env.set_input_stim_ids(g, [0, 1], f)
    else:
        with pytest.raises(pkg.ScaffoldError) as e:
            env.set_input_stim_ids(g, [0, 1], f)
        assert e.value.code == pkg.ERR_INVALID_ARG


def test_index_assignment_matches_reference(pkg):
    """NORMAL populations fill indices from the front, INPUT/OUTPUT from the back
    (init_params.cu:259-265): the reference's call order GrC, glom, PC, GoC, SC, BC,
    DCN lands as GrC 0, PC 1, GoC 2, SC 3, BC 4, DCN 5, glom 6.  Projections onto
    OUTPUT populations fill from the back (init_params.cu:309-315)."""
    env = pkg.NetworkEnvironment(1, 7, 4, 1, 0)
    Cm, El, dt_ref, Ie, Vr, Vth, gL, ch = GRANULE
    kinds = [pkg.NORMAL, pkg.INPUT, pkg.NORMAL, pkg.NORMAL, pkg.NORMAL, pkg.NORMAL, pkg.OUTPUT]
    got = [env.set_neuron_params(t, f"p{t}", 0, 10 + t, Cm, El, dt_ref, Ie, Vr, Vth, gL, k, ch)
           for t, k in enumerate(kinds)]
    assert got == [0, 6, 1, 2, 3, 4, 5]
    rp = np.zeros(11 + 1, np.int64)
    j0 = env.set_connectivity_csr(0, 10, 11, 0, 1, 0.5, 1.0, rp[:12], [])
    j1 = env.set_connectivity_csr(1, 10, 16, 0, 5, 0.5, 1.0, np.zeros(17, np.int64), [])
    j2 = env.set_connectivity_csr(2, 10, 11, 0, 1, 0.5, 1.0, rp[:12], [])
    assert (j0, j1, j2) == (0, 3, 1)
    with pytest.raises(pkg.ScaffoldError) as e:    # INPUT populations receive nothing
        env.set_connectivity_csr(3, 10, 16, 0, 6, 0.5, 1.0, np.zeros(17, np.int64), [])
    assert e.value.code == pkg.ERR_INVALID_ARG
    info = env.info()
    assert info.total_neurons == sum(10 + t for t in range(7))
    assert env.population(6).num == 11 and env.population(6).device_type == pkg.INPUT
    assert env.population(0).ref_step == 15


def _two_pop_env(pkg, n_pre=5, n_post=4, n_conn=2, n_inputs=1):
    env = pkg.NetworkEnvironment(1, 2, n_conn, n_inputs, 0)
    Cm, El, dt_ref, Ie, Vr, Vth, gL, ch = GRANULE
    a = env.set_neuron_params(0, "pre", 0, n_pre, Cm, El, dt_ref, Ie, Vr, Vth, gL, pkg.NORMAL, ch)
    b = env.set_neuron_params(1, "post", 0, n_post, Cm, El, dt_ref, Ie, Vr, Vth, gL, pkg.NORMAL, ch)
    return env, a, b


def test_dat_loader(pkg, tmp_path):
    env, a, b = _two_pop_env(pkg)
    f = tmp_path / "a_to_b.dat"
    # unsorted pairs, a blank line, a duplicate pair: rows come out grouped by post in file order
    f.write_text("3\t1\n0\t0\n\n4\t1\n3\t1\n2\t3\n")
    idx = env.set_connectivity_params(0, f, 5, 4, a, b, 0.25, 2.0)
    rp, col, val = env.projection_csr(idx)
    assert list(rp) == [0, 1, 4, 4, 5]
    assert list(col) == [0, 3, 4, 3, 2]
    assert (val == 1.0).all()
    p = env.projection(idx)
    assert p.nnz == 5 and p.max_conv == 3 and p.delay == 2 and p.w_bar == np.float32(0.25)


@pytest.mark.parametrize("text,code", [("0\t9\n", "ERR_FORMAT"), ("7\t0\n", "ERR_FORMAT"), ("0 x\n", "ERR_FORMAT"),
                                       ("-1\t0\n", "ERR_FORMAT"), (None, "ERR_IO")])
def test_dat_loader_errors(pkg, tmp_path, text, code):
    env, a, b = _two_pop_env(pkg)
    f = tmp_path / "bad.dat"
    if text is not None:
        f.write_text(text)
    with pytest.raises(pkg.ScaffoldError) as e:
        env.set_connectivity_params(0, f, 5, 4, a, b, 0.25, 2.0)
    assert e.value.code == getattr(pkg, code)
    if text and text.startswith("0\t9"):
        assert "bad.dat:1" in str(e.value)


def test_rows_beyond_the_post_population_are_dropped(pkg, tmp_path):
    """postNum larger than the post population (the reference passes
    granule_cell_num for the Golgi cells, init_params.cu:743): pairs whose post id
    lies beyond the population are dropped, as the reference never reads those
    ELL rows; ids at or beyond postNum stay a format error."""
    env, a, b = _two_pop_env(pkg)
    f = tmp_path / "wide.dat"
    f.write_text("1\t0\n2\t3\n4\t5\n0\t9\n")       # posts 5 and 9 lie beyond the 4 post cells
    idx = env.set_connectivity_params(0, f, 5, 10, a, b, 0.25, 1.0)
    rp, col, _ = env.projection_csr(idx)
    assert list(rp) == [0, 1, 1, 1, 2] and list(col) == [1, 2]
    idx = env.set_connectivity_csr(0, 5, 6, a, b, 0.25, 1.0, [0, 1, 1, 1, 2, 3, 4], [1, 2, 4, 0])
    rp, col, _ = env.projection_csr(idx)
    assert list(rp) == [0, 1, 1, 1, 2] and list(col) == [1, 2]
    env2, a2, b2 = _two_pop_env(pkg)
    f.write_text("1\t0\n0\t10\n")
    with pytest.raises(pkg.ScaffoldError) as e:
        env2.set_connectivity_params(0, f, 5, 10, a2, b2, 0.25, 1.0)
    assert e.value.code == pkg.ERR_FORMAT


def test_short_ring_with_one_step_delays_is_planned(pkg):
    """T_print = T_cpu_sim = 1 gives the reference a ring of 2 rows; with
    one-step delays a step would read the row the step before it clears, so the
    planner grows the ring to delay + 2 rows instead of accepting the race.  On
    a machine without a GPU the call gets past the planner and stops at the
    device."""
    env, a, b = _two_pop_env(pkg)
    env.set_connectivity_csr(0, 5, 4, a, b, 0.1, 1.0, [0, 1, 2, 3, 4], [0, 1, 2, 3])
    if HAS_GPU_NODE:
        pytest.skip("a GPU driver node is present (the GPU test checks the ring)")
    with pytest.raises(pkg.ScaffoldError) as e:
        env.initialize(1, 1)
    assert e.value.code in (pkg.ERR_NO_DEVICE, pkg.ERR_HIP)


def test_projection_validation(pkg):
    env, a, b = _two_pop_env(pkg)
    with pytest.raises(pkg.ScaffoldError) as e:          # sub-step delays are not exchange steps
        env.set_connectivity_csr(0, 5, 4, a, b, 0.1, 0.5, np.zeros(5, np.int64), [])
    assert e.value.code == pkg.ERR_INVALID_ARG
    with pytest.raises(pkg.ScaffoldError) as e:          # pre id out of range
        env.set_connectivity_csr(0, 5, 4, a, b, 0.1, 1.0, [0, 1, 1, 1, 1], [5])
    assert e.value.code == pkg.ERR_FORMAT
    with pytest.raises(pkg.ScaffoldError) as e:          # decreasing row_ptr
        env.set_connectivity_csr(0, 5, 4, a, b, 0.1, 1.0, [0, 2, 1, 2, 2], [0, 1])
    assert e.value.code == pkg.ERR_FORMAT
    with pytest.raises(pkg.ScaffoldError) as e:          # unknown population
        env.set_connectivity_csr(0, 5, 4, a, 7, 0.1, 1.0, np.zeros(5, np.int64), [])
    assert e.value.code == pkg.ERR_INVALID_ARG
    idx = env.set_connectivity_csr(0, 5, 4, a, b, -0.5, 3.0, [0, 2, 2, 3, 3], [1, 4, 0], [1.0, 0.5, 2.0])
    rp, col, val = env.projection_csr(idx)
    assert list(rp) == [0, 2, 2, 3, 3] and list(col) == [1, 4, 0] and list(val) == [1.0, 0.5, 2.0]
    p = env.projection(idx)
    assert p.w_bar == np.float32(0.5) and p.initial_weight == np.float32(-0.5) and p.delay == 3


def test_stimulus_registration(pkg, tmp_path):
    env = pkg.NetworkEnvironment(1, 2, 1, 2, 0)
    Cm, El, dt_ref, Ie, Vr, Vth, gL, ch = GRANULE
    g = env.set_neuron_params(0, "glom", 0, 20, 0, 0, 0, 0, 0, 0, 0, pkg.INPUT, ())
    n = env.set_neuron_params(1, "grc", 0, 8, Cm, El, dt_ref, Ie, Vr, Vth, gL, pkg.NORMAL, ch)
    f = tmp_path / "ids.dat"
    f.write_text("3\n\n5\n19\n")
    s0 = env.set_input_stim(g, f, pkg.stim_func_reference(0, 1))
    s1 = env.set_input_stim_ids(g, [1, 2, 7, 8, 9], pkg.stim_func_reference(1, 2), fanout_ptr=[0, 2, 2, 5])
    info, fp, tg = env.stimulus(s0)
    assert info.n_entries == 3 and list(tg) == [3, 5, 19] and list(fp) == [0, 1, 2, 3]
    info, fp, tg = env.stimulus(s1)
    assert info.n_entries == 3 and list(fp) == [0, 2, 2, 5] and list(tg) == [1, 2, 7, 8, 9]
    assert info.func.periodic_hz == 140.0
    with pytest.raises(pkg.ScaffoldError) as e:          # capacity num_of_inputs = 2
        env.set_input_stim_ids(g, [0], pkg.stim_func_reference(0, 1))
    assert e.value.code == pkg.ERR_CAPACITY
    env2 = pkg.NetworkEnvironment(1, 2, 1, 2, 0)
    g2 = env2.set_neuron_params(0, "glom", 0, 20, 0, 0, 0, 0, 0, 0, 0, pkg.INPUT, ())
    n2 = env2.set_neuron_params(1, "grc", 0, 8, Cm, El, dt_ref, Ie, Vr, Vth, gL, pkg.NORMAL, ch)
    for ids, pop, code in [([20], g2, "ERR_FORMAT"), ([0], n2, "ERR_INVALID_ARG")]:
        with pytest.raises(pkg.ScaffoldError) as e:
            env2.set_input_stim_ids(pop, ids, pkg.stim_func_reference(0, 1))
        assert e.value.code == getattr(pkg, code)


def test_options(pkg):
    env = new_env(pkg, 0)
    defaults = {"weight_precision": 0, "index_bits": 32, "graph": 0, "graph_steps": 25, "word_groups": -1,
                "lane_max_fanin": 48, "lane_synapses": 0, "stage_ratio": 0}
    for k, v in defaults.items():
        assert env.get_option(k) == v, k
    env.set_option("weight_precision", pkg.WEIGHTS_BF16)
    assert env.get_option("weight_precision") == pkg.WEIGHTS_BF16
    env.set_option("neuron_seed", 99)
    assert env.get_option("neuron_seed") == 99
    env.set_option("index_bits", 16)
    assert env.get_option("index_bits") == 16
    assert env.get_option("stage_spikes") == 1
    env.set_option("stage_spikes", 0)
    assert env.get_option("stage_spikes") == 0
    assert env.get_option("nontemporal") == -1
    env.set_option("nontemporal", 1)
    assert env.get_option("nontemporal") == 1
    env.set_option("stage_ratio", 32)
    assert env.get_option("stage_ratio") == 32
    env.set_option("lane_synapses", 64)
    assert env.get_option("lane_synapses") == 64
    env.set_option("word_groups", 0)
    assert env.get_option("word_groups") == 0
    env.set_option("word_groups", 64)
    assert env.get_option("word_groups") == 64
    for k, v, code in [("no_such_option", 1, "ERR_INVALID_ARG"), ("weight_precision", 7, "ERR_INVALID_ARG"),
                       ("index_bits", 8, "ERR_INVALID_ARG"), ("graph_steps", 0, "ERR_INVALID_ARG"),
                       ("nontemporal", 2, "ERR_INVALID_ARG"), ("stage_ratio", -1, "ERR_INVALID_ARG"),
                       ("stage_ratio", 1000001, "ERR_INVALID_ARG"), ("lane_synapses", -1, "ERR_INVALID_ARG"),
                       ("lane_synapses", 4097, "ERR_INVALID_ARG"), ("word_groups", -2, "ERR_INVALID_ARG"),
                       ("word_groups", 65, "ERR_INVALID_ARG")]:
        with pytest.raises(pkg.ScaffoldError) as e:
            env.set_option(k, v)
        assert e.value.code == getattr(pkg, code), k


def test_queries_before_initialization(pkg):
    env = new_env(pkg, 0)
    for call in (lambda: env.loop_n_steps(0, 25), lambda: env.state(), lambda: env.counts(0),
                 lambda: env.reset(0), lambda: env.set_state(V=np.zeros(1, np.float32)),
                 lambda: env.step_timer(1)):
        with pytest.raises(pkg.ScaffoldError) as e:
            call()
        assert e.value.code == pkg.ERR_STATE


@pytest.mark.skipif(HAS_GPU_NODE, reason="a GPU driver node is present")
def test_initialize_without_device_fails_cleanly(pkg):
    env = new_env(pkg, 0)
    with pytest.raises(pkg.ScaffoldError) as e:
        env.initialize(50, 25)
    assert e.value.code in (pkg.ERR_NO_DEVICE, pkg.ERR_HIP)
    # the description stays usable and the environment can be finalized
    assert env.info().total_neurons > 9000
    env.finalize()


def test_initialize_argument_checks(pkg):
    env = new_env(pkg, 0)
    for tp, tc in [(50, 0), (50, 30)]:
        with pytest.raises(pkg.ScaffoldError) as e:
            env.initialize(tp, tc)
        assert e.value.code == pkg.ERR_INVALID_ARG


# ------------------------------------------------------------ synthetic scaffold

def test_synthetic_scaffold_shape(pkg):
    """Sizes and fan-in statistics of the small scaffold follow the reference's
    connection rules (counts per post cell, delays, weight sign per projection)."""
    env = new_env(pkg, 0)
    info = env.info()
    names = [env.population(i).name.decode() for i in range(info.n_populations)]
    nums = {n: env.population(i).num for i, n in enumerate(names)}
    assert nums["granule_cell"] == 8800 and nums["purkinje_cell"] == 7 and nums["dcn_cell"] == 2
    assert env.population(names.index("glomerulus")).device_type == pkg.INPUT
    assert env.population(names.index("dcn_cell")).device_type == pkg.OUTPUT
    by_pair = {}
    for q in range(info.n_projections):
        p = env.projection(q)
        if p.pre < 0:
            continue
        rp, col, val = env.projection_csr(q)
        assert p.nnz == rp[-1] and 1 <= p.delay <= 4
        assert (col >= 0).all() and (col < env.population(p.pre).num).all()
        assert (val >= 0.8 - 1e-6).all() and (val <= 1.2 + 1e-6).all()
        # the rows are sorted and free of duplicates
        for j in range(0, p.post_num, max(1, p.post_num // 50)):
            r = col[rp[j]:rp[j + 1]]
            assert (np.diff(r) > 0).all()
        by_pair[(names[p.pre], names[p.post])] = (p, np.diff(rp))
    fan = by_pair[("golgi_cell", "granule_cell")][1]
    assert set(np.unique(fan)) <= {2, 3} and 2.15 < fan.mean() < 2.45
    fan = by_pair[("glomerulus", "granule_cell")][1]
    assert (fan == 4).all()
    assert by_pair[("golgi_cell", "granule_cell")][0].initial_weight < 0
    assert by_pair[("granule_cell", "purkinje_cell")][0].initial_weight > 0
    fan = by_pair[("purkinje_cell", "dcn_cell")][1]
    assert (fan == 7).all()


def test_synthetic_generator_is_deterministic(pkg):
    a, b = new_env(pkg, 0), new_env(pkg, 0)
    c = new_env(pkg, 0, seed=7)
    for q in range(a.info().n_projections):
        if a.projection(q).pre < 0:
            continue
        ra, ca, va = a.projection_csr(q)
        rb, cb, vb = b.projection_csr(q)
        assert (ra == rb).all() and (ca == cb).all() and (va == vb).all()
    diff = any(not np.array_equal(a.projection_csr(q)[1], c.projection_csr(q)[1])
               for q in range(a.info().n_projections) if a.projection(q).pre >= 0)
    assert diff


def test_write_dat_round_trip(pkg, tmp_path):
    src = new_env(pkg, 0, weight_jitter=0)
    q = next(q for q in range(src.info().n_projections) if src.projection(q).pre >= 0 and src.projection(q).nnz > 0)
    p = src.projection(q)
    f = tmp_path / "proj.dat"
    src.write_dat(q, f)
    lines = f.read_text().splitlines()
    assert len(lines) == p.nnz
    pairs = [tuple(map(int, s.split("\t"))) for s in lines]
    assert pairs == sorted(pairs, key=lambda t: (t[1], t[0]))
    env, a, b = _two_pop_env(pkg, env_pop := src.population(p.pre).num, src.population(p.post).num)
    idx = env.set_connectivity_params(0, f, env_pop, p.post_num, a, b, p.initial_weight, float(p.delay))
    r0, c0, _ = src.projection_csr(q)
    r1, c1, _ = env.projection_csr(idx)
    assert (r0 == r1).all() and (c0 == c1).all()


def test_autotune_argument_checks(pkg):
    env = new_env(pkg, 0)
    with pytest.raises(pkg.ScaffoldError) as e:
        env.autotune(0)
    assert e.value.code == pkg.ERR_INVALID_ARG
    if not HAS_GPU_NODE:
        with pytest.raises(pkg.ScaffoldError) as e:
            env.autotune(10)
        assert e.value.code in (pkg.ERR_NO_DEVICE, pkg.ERR_HIP)
        # a failed tuning leaves the options as they were
        assert env.get_option("lane_synapses") == 0 and env.get_option("lane_max_fanin") == 48


def test_plasticity_registration(pkg):
    """scaffold_set_plasticity validates the rule table before any GPU work."""
    env = new_env(pkg, 0)
    with pytest.raises(pkg.ScaffoldError) as e:   # table of size 0
        env.set_plasticity(pkg.STDP_HEBB, 0, coefficients={pkg.A2_JI: -0.01})
    assert e.value.code == pkg.ERR_CAPACITY
    env = new_env(pkg, 0, plasticities=2)
    idx = {env.projection(i).type: i for i in range(16)}   # reference connection type -> index
    pf_pc, bc_pc, goc_grc = idx[0], idx[11], idx[15]
    assert env.projection(pf_pc).post == env.projection(bc_pc).post
    bad = [dict(rule=2, target=pf_pc), dict(rule=pkg.STDP_HEBB, target=99),
           dict(rule=pkg.STDP_TEACHER, target=pf_pc, teacher=-1),
           dict(rule=pkg.STDP_TEACHER, target=pf_pc, teacher=goc_grc),   # another post population
           dict(rule=pkg.STDP_HEBB, target=pf_pc, time_window=-1),
           dict(rule=pkg.STDP_HEBB, target=pf_pc, coefficients={pkg.A0: float("nan")})]
    for kw in bad:
        with pytest.raises(pkg.ScaffoldError) as e:
            env.set_plasticity(**kw)
        assert e.value.code == pkg.ERR_INVALID_ARG, kw
    assert env.set_plasticity(pkg.STDP_TEACHER, pf_pc, bc_pc, {pkg.A1_J: 0.0, pkg.A2_JI: -0.01}) == 0
    assert env.set_plasticity(pkg.STDP_HEBB, goc_grc, coefficients={pkg.A1_I: 0.01}) == 1
    assert env.projection(pf_pc).plastic == 1 and env.projection(bc_pc).plastic == 0
    with pytest.raises(pkg.ScaffoldError) as e:
        env.set_plasticity(pkg.STDP_HEBB, goc_grc)
    assert e.value.code == pkg.ERR_CAPACITY


def test_plasticity_plan_checks(pkg):
    """The planner's plasticity checks run at initialization, before any device
    work: the window fits the one-byte counts and a Teacher rule must not
    potentiate.  A ring shorter than delay + window + 2 rows (T_print 20 gives
    40 for a window of 50 behind a delay of 5) is grown, not rejected."""
    def env_with(**rule):
        env = new_env(pkg, 0, plasticities=1)
        idx = {env.projection(i).type: i for i in range(env.info().n_projections)}
        env.set_plasticity(pkg.STDP_TEACHER, idx[0], idx[11], **rule)
        return env
    cases = [(dict(coefficients={pkg.A2_JI: -0.01}, time_window=128), 100, pkg.ERR_CAPACITY),
             (dict(coefficients={pkg.A1_J: 0.01}, time_window=50), 50, pkg.ERR_INVALID_ARG)]
    for rule, t_print, code in cases:
        env = env_with(**rule)
        with pytest.raises(pkg.ScaffoldError) as e:
            env.initialize(t_print, 10 if t_print == 20 else 25)
        assert e.value.code == code, (rule, t_print, str(e.value))
        env.finalize()
    if not HAS_GPU_NODE:  # valid rules pass the planner and stop at the device
        for t_print in (50, 20):
            env = env_with(coefficients={pkg.A1_J: 0.0, pkg.A2_JI: -0.01}, time_window=50)
            with pytest.raises(pkg.ScaffoldError) as e:
                env.initialize(t_print, 10 if t_print == 20 else 25)
            assert e.value.code in (pkg.ERR_NO_DEVICE, pkg.ERR_HIP)
            env.finalize()


def test_weight_dependent_plasticity_checks(pkg):
    """scaffold_set_plasticity_affine: a Teacher rule stays bounded only if every
    weight-dependent term decreases as w grows (a > 0), no constant term is
    positive and the potentiating slopes (times their multiplicity in a step:
    the window for A2_JI) sum to at most 1; the template's PF->PC coefficients
    (A1_J = 0.001*(1 - w), A2_JI = -0.01*w) pass.  Non-finite values are refused at
    registration."""
    def env_with(coefficients, window=50):
        env = new_env(pkg, 0, plasticities=1)
        idx = {env.projection(i).type: i for i in range(env.info().n_projections)}
        env.set_plasticity_affine(pkg.STDP_TEACHER, idx[0], idx[11], coefficients, window)
        return env
    bad = [{pkg.A1_J: (0.0, -0.001, 1.0)},                  # grows with w
           {pkg.A1_J: (0.001, 0.0, 0.0)},                   # a positive constant
           {pkg.A2_JI: (0.0, 0.05, 1.0)},                   # 50 x 0.05 > 1
           {pkg.A1_J: (0.0, 0.6, 1.0), pkg.A1_I: (0.0, 0.6, 2.0)}]
    for coefficients in bad:
        env = env_with(coefficients)
        with pytest.raises(pkg.ScaffoldError) as e:
            env.initialize(50, 25)
        assert e.value.code == pkg.ERR_INVALID_ARG, (coefficients, str(e.value))
        env.finalize()
    env = new_env(pkg, 0, plasticities=1)
    with pytest.raises(pkg.ScaffoldError) as e:
        env.set_plasticity_affine(pkg.STDP_HEBB, 0, coefficients={pkg.A0: (float("nan"), 0.0, 0.0)})
    assert e.value.code == pkg.ERR_INVALID_ARG
    env.finalize()
    if not HAS_GPU_NODE:  # the template's rule passes the planner and stops at the device
        env = env_with({pkg.A1_J: (0.0, 0.001, 1.0), pkg.A2_JI: (0.0, 0.01, 0.0)})
        with pytest.raises(pkg.ScaffoldError) as e:
            env.initialize(50, 25)
        assert e.value.code in (pkg.ERR_NO_DEVICE, pkg.ERR_HIP)
        env.finalize()
