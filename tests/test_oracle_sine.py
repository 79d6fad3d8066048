"""The oracle's sinusoidally modulated Poisson stimulus (the reference's puff_stim).

Rules followed (SimulationModule/Scaffold/UserInputFunctions.cu):
  SinusoidalOscillation :12-14  max/2 * (1 - cosf(2*M_PI*(time/sec)/osci + shift))
  puff_stim             :33-39  (0 <= time) && PoissonProcess(r, time, sine, 0)
  PoissonProcess        :4-6    r < freq/sec*DT
The cosine is the library's own (CUDA's cosf is not restated): it is pinned to
numpy's double cosine rounded to float.  The rate is pinned by values the
mathematics fixes (0 at phase 0, max/2 at phase pi/2, max at phase pi) and
decision by decision against a numpy restatement that uses numpy's cosine.
"""
import numpy as np
import pytest

from test_oracle_network import draws


def test_cosf_is_the_float_rounding_of_cos(orc):
    rng = np.random.default_rng(5)
    xs = [rng.uniform(-60, 60, 100_000), rng.uniform(0, 4000, 50_000), 10.0 ** rng.uniform(-9, 0, 5_000)]
    # float arguments next to multiples of pi/2, where cos is near 0 or near +-1
    near = np.float32(np.arange(1, 2000) * (np.pi / 2))
    for d in range(-3, 4):
        xs.append(near.view(np.int32) + np.int32(d))
    xs = np.concatenate([np.asarray(xs[0], np.float32), np.asarray(xs[1], np.float32),
                         np.asarray(xs[2], np.float32)] + [a.view(np.float32) for a in xs[3:]] +
                        [np.float32([0.0, -0.0, np.pi, -np.pi, 1e-30])])
    got = np.array([orc.cosf(float(x)) for x in xs], np.float32)
    want = np.cos(xs.astype(np.float64)).astype(np.float32)
    bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, (xs[bad[:5]], got[bad[:5]], want[bad[:5]])


def _sine_stim(orc, seed, max_hz, cycle, shift, on=0.0, off=float("inf")):
    return orc.make_stim(0, [0], seed=seed, poisson_hz=0.0, sine_max_hz=max_hz, sine_cycle_s=cycle,
                         sine_shift=shift, sine_on=on, sine_off=off)


def test_rate_at_fixed_phases(orc):
    """max 10 kHz, period 2 s: at t = 1000 ms the phase is pi and the rate is the
    full 10 kHz (threshold 1.0, so every draw fires); in the first step the rate
    is below 0.05 Hz (no entry fires); with the phase shifted by pi/2 the first
    step runs at max/2 = 5 kHz (threshold 0.5 per sub-step)."""
    st = _sine_stim(orc, 3, 10000.0, 2.0, 0.0)
    fires = lambda s, n, e: orc.lib().orc_stim_entry_fires(s, 0.0, e, n, n)
    assert all(fires(st, 1000, e) for e in range(200))
    assert not any(fires(st, 0, e) for e in range(2000))
    sh = _sine_stim(orc, 3, 10000.0, 2.0, np.pi / 2)
    # sub-step 0 at t = 0 has the threshold 0.5 exactly (cos(float(pi/2)) rounds
    # to -4.37e-8, so the rate is 5000 * (1 + 4.37e-8) = 5000 in float)
    assert np.float32(np.float32(5000.0) * np.float32(1 - np.float32(np.cos(np.float64(np.float32(np.pi / 2))))))\
        == np.float32(5000.0)
    hits = sum(fires(sh, 0, e) for e in range(2000))
    assert hits > 1990                # 1 - 0.5^10 per entry


def test_decisions_match_numpy_restatement(orc):
    """50 ms period, phase 0.7, window [3, 250) ms, peak 4 kHz (thresholds up to
    0.4): 16 entries x 300 steps, every decision against numpy, including the
    sub-steps on either side of both window edges."""
    max_hz, cycle, shift, on, off, seed = 4000.0, 0.05, 0.7, 3.0, 250.0, 17
    st = _sine_stim(orc, seed, max_hz, cycle, shift, on, off)
    f32, f64 = np.float32, np.float64
    fired = 0
    for e in range(16):
        for n in range(300):
            r = draws(orc, seed, e, n)
            want = False
            for k in range(10):
                t = f32(f64(f32(n)) + f64(f32(k)) * 0.1)
                if not (on <= f64(t) < off):
                    continue
                x = f32(2 * np.pi * (f64(t) / 1000.0) / f64(f32(cycle)) + f64(f32(shift)))
                rate = f32(f32(max_hz) / f32(2)) * f32(f32(1) - f32(np.cos(f64(x))))
                want |= f64(r[k]) < f64(rate) / 1000.0 * 0.1
            got = orc.lib().orc_stim_entry_fires(st, 0.0, e, n, n)
            assert got == want, (e, n)
            fired += got
    assert 300 < fired < 16 * 300 - 300


def test_reference_puff_rate(orc):
    """puff_stim's 3 Hz peak at t = 1000 ms: P(fire in a step) = 1 - (1 - 3e-4)^10."""
    st = _sine_stim(orc, 9, 3.0, 2.0, 0.0)
    n_e = 40_000
    hits = sum(orc.lib().orc_stim_entry_fires(st, 0.0, e, 1000, 1000) for e in range(n_e))
    p = 1 - (1 - 3e-4) ** 10
    assert abs(hits - n_e * p) < 5 * np.sqrt(n_e * p)
