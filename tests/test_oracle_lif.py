"""The oracle's neuron update against closed forms of the reference model.

The model (update_lif_ch, SimulationModule/Scaffold/Simulation_header.cu:305-381) is forward Euler with
DT = 0.1 ms on  Cm dV/dt = Ie - gL (V - El) - sum_c g_c (V - E_c),
dg_c/dt = -g_c / tau_c, synaptic input injected at the first of the 10 sub-steps
of an exchange step, a threshold Vth, reset to Vr and a refractory hold of
ref_step = (int)(dt_ref / DT) sub-steps.  The checks use closed forms of that
recursion (geometric decay, fixed points, first-passage time), not the oracle's
expression tree."""
import math

import numpy as np
import pytest

# (Cm, El, dt_ref, Ie, Vr, Vth, gL, channels as (w, E, g_bar, tau) x 2)
# -- SimulationModule/Scaffold/init_params.cu:377-528
POPS = {
    "granule": (3.0, -74.0, 1.5, 0.0, -84.0, -42.0, 1.5, (1., 0., 1.0, 0.5, 1., -85., 1.0, 10.0)),
    "purkinje": (620.0, -62.0, 0.8, 600.0, -72.0, -47.0, 7.0, (1., 0., 1.0, 0.5, 1., -85., 1.0, 1.6)),
    "golgi": (76.0, -65.0, 2.0, 36.75, -75.0, -55.0, 3.6, (1., 0., 1.0, 0.5, 1., -85., 1.0, 15.0)),
    "stellate": (14.6, -68.0, 1.6, 15.6, -78.0, -53.0, 1.0, (1., 0., 1.0, 0.64, 1., -85., 1.0, 2.0)),
    "dcn": (89.0, -59.0, 3.7, 55.8, -69.0, -48.0, 1.56, (1., 0., 1.0, 7.1, 4., -85., 1.0, 13.6)),
}


def pop(orc, name):
    Cm, El, dt_ref, Ie, Vr, Vth, gL, ch = POPS[name]
    return orc.make_pop(1, 0, Cm, El, dt_ref, Ie, Vr, Vth, gL, ch)


def test_ref_step(orc):
    # Appendix values computed from (int)((double)(float)dt_ref / 0.1); 1.5/0.1 must not
    # truncate to 14 (the float quotient would)
    want = {"granule": 15, "purkinje": 8, "golgi": 20, "stellate": 16, "dcn": 37}
    for name, r in want.items():
        assert orc.lib().orc_ref_step(POPS[name][2]) == r


@pytest.mark.parametrize("name", ["purkinje", "golgi", "stellate", "dcn"])
@pytest.mark.parametrize("f64", [False, True])
def test_free_running_interval(orc, name, f64):
    """Without input, V relaxes geometrically from Vr towards V_inf = El + Ie/gL with
    factor (1 - DT gL / Cm) per sub-step; the spike comes at the first sub-step j
    with V_j > Vth after the ref_step-long hold.  Interval = ref_step + j*."""
    Cm, El, dt_ref, Ie, Vr, Vth, gL, _ = POPS[name]
    p = pop(orc, name)
    a = 0.1 * gL / Cm
    vinf = El + Ie / gL
    jstar = math.ceil(math.log((Vth - vinf) / (Vr - vinf)) / math.log(1.0 - a))
    isi = orc.lib().orc_ref_step(dt_ref) + jstar
    u, g, r = Vr, [0.0, 0.0], 0
    spikes = []
    for step in range(6 * isi // 10 + 20):
        s, u, g, r = orc.lif_step(p, u, g, r, 0.0, 0.0, f64=f64)
        if s:
            spikes.append(step)
    # spikes are separated by isi sub-steps; at exchange-step resolution the gaps
    # are floor/ceil of isi/10
    gaps = np.diff(spikes)
    assert len(spikes) >= 4
    assert set(gaps) <= {isi // 10, -(-isi // 10)}
    span = (spikes[-1] - spikes[0]) * 10
    assert abs(span / (len(spikes) - 1) - isi) <= 10.0 / (len(spikes) - 1) + 1e-9


def test_free_running_intervals(orc):
    # free-running intervals in sub-steps quoted for the reference model
    want = {"purkinje": 276, "golgi": 984, "stellate": 563, "dcn": 387}
    for name, isi in want.items():
        Cm, El, dt_ref, Ie, Vr, Vth, gL, _ = POPS[name]
        a = 0.1 * gL / Cm
        vinf = El + Ie / gL
        j = math.ceil(math.log((Vth - vinf) / (Vr - vinf)) / math.log(1.0 - a))
        assert orc.lib().orc_ref_step(dt_ref) + j == isi


def test_granule_fixed_point(orc):
    """Granule cells have Ie = 0, so V -> El from below and never fire."""
    p = pop(orc, "granule")
    u, g, r = -84.0, [0.0, 0.0], 0
    for step in range(400):
        s, u, g, r = orc.lif_step(p, u, g, r, 0.0, 0.0)
        assert s == 0
    assert abs(u - (-74.0)) < 1e-3
    # the geometric approach: after 10 sub-steps the distance shrinks by (1 - a)^10
    u, g, r = -84.0, [0.0, 0.0], 0
    _, u1, _, _ = orc.lif_step(p, u, g, r, 0.0, 0.0)
    a = 0.1 * 1.5 / 3.0
    assert abs((u1 + 74.0) - (-10.0) * (1 - a) ** 10) < 1e-4


@pytest.mark.parametrize("f64", [False, True])
def test_conductance_injection_and_decay(orc, f64):
    """dg enters at sub-step 0 scaled by g_bar*w, then every sub-step multiplies
    by (1 - DT/tau): after one exchange step g = g_bar w dg (1 - DT/tau)^9."""
    p = pop(orc, "granule")
    _, _, g, _ = orc.lif_step(p, -84.0, [0.0, 0.0], 0, 1.0, 0.0, f64=f64)
    assert abs(g[0] - 0.8 ** 9) < 1e-6 and g[1] == 0.0
    _, _, g2, _ = orc.lif_step(p, -84.0, g, 0, 0.0, 0.0, f64=f64)
    assert abs(g2[0] - 0.8 ** 19) < 1e-7
    _, _, g, _ = orc.lif_step(p, -84.0, [0.0, 0.0], 0, 0.0, 2.0, f64=f64)
    assert g[0] == 0.0 and abs(g[1] - 2.0 * 0.99 ** 9) < 1e-6
    # DCN inhibitory channel has w = 4: injection g_bar * w * dg
    d = pop(orc, "dcn")
    _, _, g, _ = orc.lif_step(d, -69.0, [0.0, 0.0], 0, 0.0, 0.5, f64=f64)
    assert abs(g[1] - 4 * 0.5 * (1 - 0.1 / 13.6) ** 9) < 1e-5


def test_excitation_and_inhibition_move_v(orc):
    p = pop(orc, "granule")
    base = orc.lif_step(p, -74.0, [0.0, 0.0], 0, 0.0, 0.0)[1]
    up = orc.lif_step(p, -74.0, [0.0, 0.0], 0, 3.0, 0.0)[1]
    down = orc.lif_step(p, -74.0, [0.0, 0.0], 0, 0.0, 3.0)[1]
    assert down < base < up


def test_refractory_hold(orc):
    """A neuron above threshold fires at sub-step 0 and is held at Vr for ref_step
    sub-steps: after the step V == Vr exactly and ref_step - 9 sub-steps remain."""
    p = pop(orc, "granule")
    s, u, g, r = orc.lif_step(p, 0.0, [0.0, 0.0], 0, 0.0, 0.0)
    assert s == 1 and u == np.float32(-84.0) and r == 15 - 9
    # the next step releases after 6 sub-steps and integrates the remaining 4:
    # V = El + (Vr - El) (1 - a)^4 ... one free sub-step happens at the release itself
    s, u, g, r = orc.lif_step(p, u, g, r, 0.0, 0.0)
    a = 0.1 * 1.5 / 3.0
    assert s == 0 and r == 0
    assert abs((u + 74.0) - (-10.0) * (1 - a) ** 5) < 1e-4


def test_hand_computed_first_substep(orc):
    """u = -60, g_exc = 1, granule cell: Isyn = -60, du = (0.1/3)(0 - 1.5*14 + 60) = 1.3;
    after 10 sub-steps the recursion is V_{k+1} = V_k + (0.1/3)(-1.5 (V_k + 74) - g_k V_k)
    with g_k = 0.8^k, evaluated here in double from the model equations."""
    p = pop(orc, "granule")
    V, gk = -60.0, 1.0
    for k in range(10):
        V = V + (0.1 / 3.0) * (0.0 - 1.5 * (V + 74.0) - gk * (V - 0.0))
        gk *= 0.8
    _, u, g, _ = orc.lif_step(p, -60.0, [1.0, 0.0], 0, 0.0, 0.0)
    assert abs(u - V) < 1e-4
    assert abs(g[0] - 0.8 ** 10) < 1e-6
