"""The oracle's plasticity rules against counts taken by hand from the rule text.

Every neuron is driven (driven mode with no simulated neuron), so the spike trains
are chosen by the test.  Rules followed (SimulationModule/Scaffold/...):
  Hebb_plasticity     SynapticPlasticity.cu:38-76   dw = A0 + A1_J [pre at n-d]
                      + A1_I [post at n] + [post at n] * sum_{s<W} A2_JI [pre at n-d-s];
                      w = clamp(w + dw, 0, 1)
  Teacher_plasticity  SynapticPlasticity.cu:78-121  as Hebb with the teacher (first
                      entry of the post neuron's row of the teacher projection) at
                      n - t_delay in place of the post neuron; w = max(w + dw, 0)
  invoke_stdp_plasticity GPUScaffoldSimulation.cu:754 (after the neuron update)
With dyadic coefficients every float sum is exact, so the expected multipliers
follow from spike counts alone, independently of the summation order; one test
pins that order separately with the reference's -0.01.
"""
import ctypes

import numpy as np
import pytest

N_PRE, N_POST, N_T = 5, 3, 3
# target 0 -> 1 (delay 2); teacher 2 -> 1 (delay 1): post 0 <- [1, 2] (teacher 1),
# post 1 <- [] (no teacher), post 2 <- [0] (teacher 0)
T_RP = np.array([0, 3, 4, 6], np.int64)
T_COL = np.array([0, 1, 2, 3, 1, 4], np.int32)
TE_RP = np.array([0, 2, 2, 3], np.int64)
TE_COL = np.array([1, 2, 0], np.int32)
TEACHER_OF = {0: 1, 1: None, 2: 0}


def build(orc, R, val, delay=2, t_delay=1):
    pops = [orc.make_pop(N_PRE, 1), orc.make_pop(N_POST, 0, 3.0, -74.0, 1.5, 0.0, -84.0, -42.0, 1.5,
                                                (1., 0., 1.0, 0.5)),
            orc.make_pop(N_T, 1)]
    keep = [T_RP, T_COL, TE_RP, TE_COL, val]
    projs = []
    for pre, rp, col, v, d in ((0, T_RP, T_COL, val, delay), (2, TE_RP, TE_COL, None, t_delay)):
        o = orc.OrcProj()
        o.pre, o.post, o.w_bar, o.delay = pre, 1, 0.5, d
        o.row_ptr = rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        o.col = col.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        o.val = None if v is None else v.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        projs.append(o)
    sim = orc.OracleSim(pops, projs, [], R, 1, keep=keep)
    sim.initialize()
    sim.set_active(np.zeros(sim.total, np.uint8))   # every neuron driven
    return sim


def trains(seed, steps, p=0.3):
    rng = np.random.default_rng(seed)
    return (rng.random((steps, N_PRE + N_POST + N_T)) < p).astype(np.uint8)


def at(spk, n, i):
    return int(spk[n, i]) if n >= 0 else 0


def expected_step(w, spk, n, rule, coef, window, delay, t_delay):
    """Multipliers after step n from counts (dyadic coefficients: exact)."""
    a0, a1i, a1j, a2 = (coef.get(k, 0.0) for k in range(4))
    out = w.copy()
    for j in range(N_POST):
        if rule == 0:
            s_post = at(spk, n, N_PRE + j)
        else:
            t = TEACHER_OF[j]
            s_post = 0 if t is None else at(spk, n - t_delay, N_PRE + N_POST + t)
        for e in range(T_RP[j], T_RP[j + 1]):
            pre = int(T_COL[e])
            cnt = sum(at(spk, n - delay - s, pre) for s in range(window))
            dw = a0 + a1j * at(spk, n - delay, pre) + a1i * s_post + a2 * s_post * cnt
            x = float(w[e]) + dw
            x = min(max(x, 0.0), 1.0) if rule == 0 else max(x, 0.0)
            out[e] = np.float32(x)
    return out


@pytest.mark.parametrize("rule", [0, 1])
def test_rules_match_spike_counts(orc, rule):
    """60 steps of random trains, window 10, ring of 16 rows (wraps): after every
    step each multiplier equals the count formula (both clamps are reached)."""
    coef = {0: -1 / 1024, 1: 1 / 128, 2: 1 / 256, 3: -1 / 64} if rule == 0 else \
           {0: -1 / 1024, 2: -1 / 256, 3: -1 / 64}
    val = np.array([0.5, 0.25, 1.0, 0.75, 0.125, 0.9375], np.float32)
    sim = build(orc, 16, val)
    sim.add_plasticity(rule, 0, 1 if rule == 1 else -1, coef, window=10, grid_exp=-40)
    spk = trains(3 + rule, 60)
    w = val.copy()
    hit_lo = hit_hi = False
    for n in range(60):
        sim.step_driven(n, spk[n])
        w = expected_step(w, spk, n, rule, coef, 10, 2, 1)
        got = sim.multipliers(0)
        assert np.array_equal(got.view(np.uint32), w.view(np.uint32)), (n, got, w)
        hit_lo |= bool((w == 0).any())
        hit_hi |= bool((w == 1).any()) and rule == 0
    assert hit_lo and (hit_hi or rule == 1)


def test_window_edges(orc):
    """Hebb, A2_JI only, window 4, delay 2: a post spike at n counts pre spikes at
    n-2 .. n-5 and not at n-1 or n-6."""
    val = np.ones(6, np.float32)
    sim = build(orc, 16, val)
    sim.add_plasticity(0, 0, -1, {3: -1 / 64}, window=4, grid_exp=-40)
    spk = np.zeros((12, N_PRE + N_POST + N_T), np.uint8)
    pre = 3                      # the only entry of post 1 (e = 3)
    for t in (3, 4, 5, 6, 7, 9):  # relative to the post spike at n = 10: n-7, n-6, n-5, n-4, n-3, n-1
        spk[t, pre] = 1
    spk[10, N_PRE + 1] = 1
    for n in range(12):
        sim.step_driven(n, spk[n])
    # counted: n-2..n-5 = steps 8, 7, 6, 5 -> spikes at 5, 6, 7 (three)
    assert sim.multipliers(0)[3] == np.float32(1 - 3 / 64)


def test_teacher_delay_and_empty_row(orc):
    """Teacher, A1_I only: the teacher is read at n - t_delay (t_delay 3 here); a
    post neuron with an empty teacher row never changes."""
    val = np.full(6, 0.5, np.float32)
    sim = build(orc, 32, val, t_delay=3)
    sim.add_plasticity(1, 0, 1, {1: -1 / 8}, window=5, grid_exp=-40)
    spk = np.zeros((10, N_PRE + N_POST + N_T), np.uint8)
    spk[2, N_PRE + N_POST + 1] = 1   # teacher of post 0 fires at step 2 -> applies at step 5
    spk[2, N_PRE + N_POST + 0] = 1   # teacher of post 2
    spk[:, N_PRE + 1] = 1            # post 1 fires every step: irrelevant for Teacher
    for n in range(10):
        sim.step_driven(n, spk[n])
        w = sim.multipliers(0)
        want0 = 0.375 if n >= 5 else 0.5
        assert (w[0:3] == np.float32(want0)).all() and w[3] == np.float32(0.5), (n, w)
        assert (w[4:6] == np.float32(want0)).all()


def test_float_order_of_the_reference(orc):
    """A2_JI = -0.01 three times: dw is summed in float first, then added to w
    (SynapticPlasticity.cu:69-72).  Adding each term to w directly rounds
    differently, so the test tells the two orders apart."""
    val = np.array([0.3, 1, 1, 1, 1, 1], np.float32)
    sim = build(orc, 64, val)
    sim.add_plasticity(0, 0, -1, {3: -0.01}, window=10, grid_exp=-60)
    spk = np.zeros((12, N_PRE + N_POST + N_T), np.uint8)
    spk[[2, 4, 6], 0] = 1            # pre 0 (entry 0, post 0) fires three times
    spk[10, N_PRE + 0] = 1           # post 0 fires at 10: window n-2 .. n-11 covers 8 .. -1
    for n in range(12):
        sim.step_driven(n, spk[n])
    f = np.float32
    dw = f(f(f(f(0) + f(-0.01)) + f(-0.01)) + f(-0.01))
    summed_first = f(f(0.3) + dw)
    one_by_one = f(f(f(f(0.3) + f(-0.01)) + f(-0.01)) + f(-0.01))
    assert summed_first != one_by_one
    assert sim.multipliers(0)[0] == summed_first


def test_grid_rounding(orc):
    """Updated multipliers are rounded to multiples of 2^grid_exp (ties to even)."""
    for grid, want in ((-10, 0.5), (-12, 0.5 - 2 ** -12), (-11, 0.5)):
        val = np.full(6, 0.5, np.float32)
        sim = build(orc, 16, val)
        sim.add_plasticity(0, 0, -1, {0: -2 ** -12}, window=1, grid_exp=grid)
        sim.step_driven(0, np.zeros(N_PRE + N_POST + N_T, np.uint8))
        assert sim.multipliers(0)[0] == np.float32(want), grid


def template_a1j(w):
    """PF_PC_A1_J of the code-generation template (convertHDF52dat/base_file/
    SynapticPlasticity.cu:14-16): 0.001*(1.0 - w), a double expression returned as
    CTYPE (float)."""
    return np.float32(0.001 * (1.0 - np.float64(w)))


def template_a2(w):
    """PF_PC_A2_TI of the template (:17-19): -0.01*w, returned as float."""
    return np.float32(-0.01 * np.float64(w))


def hebb_a1i(w):
    """A weight-dependent A1_I with a constant part: -0.002 + 0.004*(0.5 - w)."""
    return np.float32(-0.002 + 0.004 * (0.5 - np.float64(w)))


def expected_step_literal(w, spk, n, rule, window, delay, t_delay):
    """One update written out as the reference's kernel does it: dw is a float
    sum in the order A0, A1_J, A1_I, then one A2_JI per spike of the window, each
    coefficient the function value at the synapse's multiplier before the
    update; then the rule's clamp."""
    out = w.copy()
    for j in range(N_POST):
        if rule == 0:
            s_post = at(spk, n, N_PRE + j)
        else:
            t = TEACHER_OF[j]
            s_post = 0 if t is None else at(spk, n - t_delay, N_PRE + N_POST + t)
        for e in range(T_RP[j], T_RP[j + 1]):
            pre = int(T_COL[e])
            w0 = w[e]
            dw = np.float32(0.0)
            if at(spk, n - delay, pre):
                dw = np.float32(dw + template_a1j(w0))
            if rule == 0 and s_post:
                dw = np.float32(dw + hebb_a1i(w0))
            if s_post:
                for s in range(window):
                    if at(spk, n - delay - s, pre):
                        dw = np.float32(dw + template_a2(w0))
            x = np.float32(w0 + dw)
            x = np.float32(min(max(x, 0.0), 1.0)) if rule == 0 else np.float32(max(x, 0.0))
            out[e] = x
    return out


@pytest.mark.parametrize("rule", [0, 1])
def test_weight_dependent_coefficients_match_template(orc, rule):
    """Weight-dependent coefficients (c, a, b) -> (float)(c + a*(b - (double)w)):
    the template's PF_PC_A1_J = 0.001*(1.0 - w) as (0, 0.001, 1.0) and
    PF_PC_A2_TI = -0.01*w as (0, 0.01, 0.0), plus (Hebb) an A1_I with a constant
    part, against the functions written out literally in float64/float32; 80
    steps of random trains, window 10, ring of 16 rows."""
    coef = {2: (0.0, 0.001, 1.0), 3: (0.0, 0.01, 0.0)}
    if rule == 0:
        coef[1] = (-0.002, 0.004, 0.5)
    val = np.array([0.5, 0.3, 0.9, 0.7, 0.1, 0.95], np.float32)
    sim = build(orc, 16, val)
    sim.add_plasticity_affine(rule, 0, 1 if rule == 1 else -1, coef, window=10, grid_exp=-60)
    spk = trains(11 + rule, 80, p=0.35)
    w = val.copy()
    for n in range(80):
        sim.step_driven(n, spk[n])
        w = expected_step_literal(w, spk, n, rule, 10, 2, 1)
        got = sim.multipliers(0)
        assert np.array_equal(got.view(np.uint32), w.view(np.uint32)), (n, got, w)
    assert not np.array_equal(w, val)


def test_weight_dependent_potentiation_converges_to_its_fixed_point(orc):
    """Teacher rule with only A1_J = 0.25*(1 - w) and a presynaptic spike every
    step: w_{k+1} = w_k + (float)(0.25*(1 - w_k)) rises towards 1 and never above
    it (the bound the planner relies on); the sequence equals the recursion
    computed here."""
    val = np.array([0.0, 0.5, 0.75, 0.99, 0.125, 1.0], np.float32)
    sim = build(orc, 16, val)
    sim.add_plasticity_affine(1, 0, 1, {2: (0.0, 0.25, 1.0)}, window=10, grid_exp=-60)
    spk = np.zeros((120, N_PRE + N_POST + N_T), np.uint8)
    spk[:, :N_PRE] = 1
    w = val.copy()
    for n in range(120):
        sim.step_driven(n, spk[n])
        if n >= 2:   # the pre spike of step n - delay exists from step 2 on
            w = np.array([np.float32(x + np.float32(0.25 * (1.0 - np.float64(x)))) for x in w], np.float32)
        got = sim.multipliers(0)
        assert np.array_equal(got.view(np.uint32), w.view(np.uint32)), n
        assert (got <= 1.0).all()
    assert (w > 0.999).all()
