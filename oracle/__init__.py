"""TEST INFRASTRUCTURE -- ctypes binding of the CPU oracle (oracle/scaffold_oracle.c).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only to check the engine's results.  The product never imports it.

OracleSim.from_env() copies an environment's description (populations,
projections as CSR with the multipliers the engine applies, stimuli) through the
engine's query API and builds the same network in the oracle.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


class OrcPop(ctypes.Structure):
    _fields_ = [("num", ctypes.c_int), ("kind", ctypes.c_int),
                ("Cm", ctypes.c_float), ("El", ctypes.c_float), ("dt_ref", ctypes.c_float), ("Ie", ctypes.c_float),
                ("Vr", ctypes.c_float), ("Vth", ctypes.c_float), ("gL", ctypes.c_float),
                ("c_num", ctypes.c_int), ("w", ctypes.c_float * 2), ("E", ctypes.c_float * 2),
                ("g_bar", ctypes.c_float * 2), ("tau", ctypes.c_float * 2)]


class OrcProj(ctypes.Structure):
    _fields_ = [("pre", ctypes.c_int), ("post", ctypes.c_int), ("w_bar", ctypes.c_float), ("delay", ctypes.c_int),
                ("row_ptr", ctypes.POINTER(ctypes.c_int64)), ("col", ctypes.POINTER(ctypes.c_int32)),
                ("val", ctypes.POINTER(ctypes.c_float)), ("sum_scale", ctypes.c_float)]


class OrcStim(ctypes.Structure):
    _fields_ = [("pop", ctypes.c_int),
                ("poisson_hz", ctypes.c_float), ("poisson_on", ctypes.c_double), ("poisson_off", ctypes.c_double),
                ("burst_hz", ctypes.c_float), ("burst_on", ctypes.c_double), ("burst_off", ctypes.c_double),
                ("periodic_hz", ctypes.c_float), ("periodic_start", ctypes.c_float),
                ("periodic_on", ctypes.c_double), ("periodic_off", ctypes.c_double),
                ("periodic_off_inclusive", ctypes.c_int), ("poisson_hz_from_freq", ctypes.c_int),
                ("sine_max_hz", ctypes.c_float), ("sine_cycle_s", ctypes.c_float), ("sine_shift", ctypes.c_float),
                ("sine_on", ctypes.c_double), ("sine_off", ctypes.c_double),
                ("seed", ctypes.c_uint64), ("n_entries", ctypes.c_int),
                ("fanout_ptr", ctypes.POINTER(ctypes.c_int64)), ("targets", ctypes.POINTER(ctypes.c_uint32))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        u32, u64, i32, i64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int64
        L.orc_philox4x32_10.argtypes = [P(u32), P(u32), P(u32)]
        L.orc_philox4x32_10.restype = None
        L.orc_philox_u32.argtypes = [u64, u64, u64]
        L.orc_philox_u32.restype = u32
        L.orc_uniform.argtypes = [u32]
        L.orc_uniform.restype = ctypes.c_float
        L.orc_cosf.argtypes = [ctypes.c_float]
        L.orc_cosf.restype = ctypes.c_float
        L.orc_create.argtypes = [ctypes.c_int, P(OrcPop), ctypes.c_int, P(OrcProj), ctypes.c_int, P(OrcStim),
                                 ctypes.c_int, u64, ctypes.c_int, ctypes.c_int]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_total_neurons.argtypes = [ctypes.c_void_p]
        L.orc_base_id.argtypes = [ctypes.c_void_p, ctypes.c_int]
        for f in ("orc_init_params", "orc_reset"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
            getattr(L, f).restype = None
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_step.restype = None
        L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_run.restype = None
        L.orc_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_threads.restype = None
        L.orc_set_active.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.orc_set_active.restype = None
        L.orc_step_driven.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.orc_step_driven.restype = None
        L.orc_set_freq.argtypes = [ctypes.c_void_p, P(ctypes.c_float)]
        L.orc_set_freq.restype = None
        L.orc_get_state.argtypes = [ctypes.c_void_p, P(ctypes.c_float), P(ctypes.c_float), P(i32)]
        L.orc_get_state.restype = None
        L.orc_set_state.argtypes = [ctypes.c_void_p, P(ctypes.c_float), P(ctypes.c_float), P(i32)]
        L.orc_set_state.restype = None
        L.orc_get_state_f64.argtypes = [ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_double), P(i32)]
        L.orc_get_state_f64.restype = None
        L.orc_spike_row.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_spike_row.restype = P(ctypes.c_uint8)
        L.orc_steps_since_init.argtypes = [ctypes.c_void_p]
        L.orc_steps_since_init.restype = i64
        L.orc_lif_step_f32.argtypes = [P(OrcPop), P(ctypes.c_float), P(ctypes.c_float), P(i32), ctypes.c_float,
                                       ctypes.c_float]
        L.orc_lif_step_f64.argtypes = [P(OrcPop), P(ctypes.c_double), P(ctypes.c_double), P(i32), ctypes.c_double,
                                       ctypes.c_double]
        L.orc_ref_step.argtypes = [ctypes.c_float]
        L.orc_stim_entry_fires.argtypes = [P(OrcStim), ctypes.c_float, ctypes.c_int, ctypes.c_int, i64]
        L.orc_add_plasticity.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_float),
                                         ctypes.c_uint, ctypes.c_int, ctypes.c_int]
        L.orc_add_plasticity.restype = ctypes.c_int
        L.orc_set_plasticity_affine.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, P(ctypes.c_double),
                                                P(ctypes.c_double), P(ctypes.c_double)]
        L.orc_set_plasticity_affine.restype = ctypes.c_int
        L.orc_get_multipliers.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_float)]
        L.orc_get_multipliers.restype = None
        _lib = L
    return _lib


def philox_block(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def philox_u32(seed, subsequence, offset):
    return lib().orc_philox_u32(seed, subsequence, offset)


def uniform(v):
    return lib().orc_uniform(v)


def cosf(x):
    return lib().orc_cosf(x)


def make_pop(num, kind, Cm=0., El=0., dt_ref=0., Ie=0., Vr=0., Vth=0., gL=0., channels=()):
    p = OrcPop()
    p.num, p.kind = num, kind
    p.Cm, p.El, p.dt_ref, p.Ie, p.Vr, p.Vth, p.gL = Cm, El, dt_ref, Ie, Vr, Vth, gL
    ch = list(channels)
    p.c_num = len(ch) // 4
    for c in range(p.c_num):
        p.w[c], p.E[c], p.g_bar[c], p.tau[c] = ch[4 * c:4 * c + 4]
    return p


def make_stim(pop, targets, fanout_ptr=None, seed=0, poisson_hz=0., poisson_on=0., poisson_off=float("inf"),
              burst_hz=0., burst_on=0., burst_off=0., periodic_hz=0., periodic_start=0., periodic_on=0.,
              periodic_off=0., periodic_off_inclusive=0, poisson_hz_from_freq=0, sine_max_hz=0., sine_cycle_s=0.,
              sine_shift=0., sine_on=0., sine_off=0.):
    s = OrcStim()
    s.pop = pop
    s.poisson_hz, s.poisson_on, s.poisson_off = poisson_hz, poisson_on, poisson_off
    s.burst_hz, s.burst_on, s.burst_off = burst_hz, burst_on, burst_off
    s.periodic_hz, s.periodic_start, s.periodic_on, s.periodic_off = periodic_hz, periodic_start, periodic_on, periodic_off
    s.periodic_off_inclusive, s.poisson_hz_from_freq, s.seed = periodic_off_inclusive, poisson_hz_from_freq, seed
    s.sine_max_hz, s.sine_cycle_s, s.sine_shift = sine_max_hz, sine_cycle_s, sine_shift
    s.sine_on, s.sine_off = sine_on, sine_off
    s._keep = []
    tg = np.ascontiguousarray(targets, dtype=np.uint32)
    s._keep.append(tg)
    s.targets = tg.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    if fanout_ptr is None:
        s.n_entries = tg.size
        s.fanout_ptr = None
    else:
        fp = np.ascontiguousarray(fanout_ptr, dtype=np.int64)
        s._keep.append(fp)
        s.n_entries = fp.size - 1
        s.fanout_ptr = fp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    return s


def lif_step(pop, u, g, refr, dge, dgi, f64=False):
    """One exchange step of a single neuron; returns (spike, u, g, refr)."""
    ct = ctypes.c_double if f64 else ctypes.c_float
    uu = ct(u)
    gg = (ct * 2)(*g)
    rr = ctypes.c_int32(refr)
    fn = lib().orc_lif_step_f64 if f64 else lib().orc_lif_step_f32
    s = fn(ctypes.byref(pop), ctypes.byref(uu), gg, ctypes.byref(rr), dge, dgi)
    return s, uu.value, [gg[0], gg[1]], rr.value


class OracleSim:
    """The oracle simulation of one network (FP32 reference semantics or FP64)."""

    def __init__(self, pops, projs, stims, ring_rows, neuron_seed, precision=0, sum_mode=0, keep=()):
        L = lib()
        self._keep = list(keep)
        self._pops = (OrcPop * max(len(pops), 1))(*pops)
        self._projs = (OrcProj * max(len(projs), 1))(*projs)
        self._stims = (OrcStim * max(len(stims), 1))(*stims)
        self._keep += [s._keep for s in stims if hasattr(s, "_keep")]
        self.n_pops = len(pops)
        self.ring_rows = ring_rows
        self.precision = precision
        self._h = L.orc_create(len(pops), self._pops, len(projs), self._projs, len(stims), self._stims, ring_rows,
                               neuron_seed, precision, sum_mode)
        self.total = L.orc_total_neurons(self._h)
        self.base = [L.orc_base_id(self._h, p) for p in range(len(pops))]

    def __del__(self):
        try:
            lib().orc_destroy(self._h)
        except Exception:
            pass

    @classmethod
    def from_env(cls, env, precision=0, sum_mode=0, neuron_seed=None, ring_rows=None):
        info = env.info()
        pops, projs, stims, keep = [], [], [], []
        for i in range(info.n_populations):
            p = env.population(i)
            ch = []
            for c in range(p.c_num):
                ch += [p.w[c], p.E[c], p.g_bar[c], p.tau[c]]
            pops.append(make_pop(p.num, p.device_type, p.Cm, p.El, p.dt_ref, p.Ie, p.Vr, p.Vth, p.gL, ch))
        for i in range(info.n_projections):
            pi = env.projection(i)
            if pi.pre < 0:
                continue
            rp, col, val = env.projection_csr(i)
            rp = np.ascontiguousarray(rp)
            col = np.ascontiguousarray(col)
            unit = val.size == 0 or bool(np.all(val == 1.0))
            val = None if unit else np.ascontiguousarray(val)
            keep += [rp, col, val]
            o = OrcProj()
            o.pre, o.post, o.w_bar, o.delay = pi.pre, pi.post, pi.initial_weight, pi.delay
            o.sum_scale = pi.sum_scale
            o.row_ptr = rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
            o.col = col.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
            o.val = None if val is None else val.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            projs.append(o)
        for i in range(info.n_stimuli):
            s, fp, tg = env.stimulus(i)
            f = s.func
            stims.append(make_stim(s.pop, tg, fp, seed=f.seed, poisson_hz=f.poisson_hz, poisson_on=f.poisson_on,
                                   poisson_off=f.poisson_off, burst_hz=f.burst_hz, burst_on=f.burst_on,
                                   burst_off=f.burst_off, periodic_hz=f.periodic_hz, periodic_start=f.periodic_start,
                                   periodic_on=f.periodic_on, periodic_off=f.periodic_off,
                                   periodic_off_inclusive=f.periodic_off_inclusive,
                                   poisson_hz_from_freq=f.poisson_hz_from_freq, sine_max_hz=f.sine_max_hz,
                                   sine_cycle_s=f.sine_cycle_s, sine_shift=f.sine_shift, sine_on=f.sine_on,
                                   sine_off=f.sine_off))
        seed = env.get_option("neuron_seed") if neuron_seed is None else neuron_seed
        R = ring_rows if ring_rows is not None else info.ring_rows
        return cls(pops, projs, stims, R, seed, precision, sum_mode, keep)

    def init_params(self):
        lib().orc_init_params(self._h)

    def reset(self):
        lib().orc_reset(self._h)

    def initialize(self):
        """Initialization draw + the trial-0 reset, as InitializeNetworkEnv does."""
        self.init_params()
        self.reset()

    def step(self, n):
        lib().orc_step(self._h, n)

    def run(self, n0, k):
        lib().orc_run(self._h, n0, k)

    def set_threads(self, n):
        lib().orc_set_threads(self._h, n)

    def set_active(self, active):
        """Driven mode: simulate only the neurons with active[id] != 0 (None: all)."""
        self._active = None if active is None else np.ascontiguousarray(active, np.uint8)
        if self._active is not None:
            assert self._active.shape == (self.total,)
        lib().orc_set_active(self._h, None if self._active is None else self._active.ctypes.data)

    def step_driven(self, n, ext):
        """One step in driven mode; ext: spikes of step n of every neuron (bytes)."""
        e = np.ascontiguousarray(ext, np.uint8)
        assert e.shape == (self.total,)
        lib().orc_step_driven(self._h, n, e.ctypes.data)

    def set_freq(self, freqs):
        f = np.ascontiguousarray(freqs, dtype=np.float32)
        lib().orc_set_freq(self._h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))

    def state(self):
        V = np.zeros(self.total, np.float32)
        g = np.zeros(2 * self.total, np.float32)
        r = np.zeros(self.total, np.int32)
        P = ctypes.POINTER
        lib().orc_get_state(self._h, V.ctypes.data_as(P(ctypes.c_float)), g.ctypes.data_as(P(ctypes.c_float)),
                            r.ctypes.data_as(P(ctypes.c_int32)))
        return V, g, r

    def state_f64(self):
        V = np.zeros(self.total, np.float64)
        g = np.zeros(2 * self.total, np.float64)
        r = np.zeros(self.total, np.int32)
        P = ctypes.POINTER
        lib().orc_get_state_f64(self._h, V.ctypes.data_as(P(ctypes.c_double)), g.ctypes.data_as(P(ctypes.c_double)),
                                r.ctypes.data_as(P(ctypes.c_int32)))
        return V, g, r

    def set_state(self, V=None, g=None, refr=None):
        P = ctypes.POINTER
        V = None if V is None else np.ascontiguousarray(V, np.float32)
        g = None if g is None else np.ascontiguousarray(g, np.float32)
        refr = None if refr is None else np.ascontiguousarray(refr, np.int32)
        lib().orc_set_state(self._h, None if V is None else V.ctypes.data_as(P(ctypes.c_float)),
                            None if g is None else g.ctypes.data_as(P(ctypes.c_float)),
                            None if refr is None else refr.ctypes.data_as(P(ctypes.c_int32)))

    def add_plasticity(self, rule, target, teacher=-1, coefficients=None, window=50, grid_exp=-126):
        """A plasticity rule (SynapticPlasticity.cu:38-121); `coefficients` maps slots
        0..3 (A0, A1_I, A1_J, A2_JI) to constants, absent slots are unset."""
        c = (ctypes.c_float * 4)()
        present = 0
        for slot, v in (coefficients or {}).items():
            if slot < 4:
                c[slot] = v
                present |= 1 << slot
        if lib().orc_add_plasticity(self._h, rule, target, teacher, c, present, window, grid_exp) != 0:
            raise ValueError("bad plasticity rule")
        self._n_plast = getattr(self, "_n_plast", 0) + 1

    def add_plasticity_affine(self, rule, target, teacher=-1, coefficients=None, window=50, grid_exp=-126):
        """A plasticity rule whose slot c (A0, A1_I, A1_J, A2_JI) maps to (c0, a, b):
        the coefficient is the float value of c0 + a * (b - (double)w) for the
        synapse's multiplier w before the update; a = 0 is the constant c0."""
        self.add_plasticity(rule, target, teacher, {k: (v[0] if v[1] == 0.0 else 0.0)
                                                    for k, v in (coefficients or {}).items()}, window, grid_exp)
        ac, aa, ab = (ctypes.c_double * 4)(), (ctypes.c_double * 4)(), (ctypes.c_double * 4)()
        affine = 0
        for slot, (c0, a, b) in (coefficients or {}).items():
            if slot < 4 and a != 0.0:
                ac[slot], aa[slot], ab[slot] = c0, a, b
                affine |= 1 << slot
        if lib().orc_set_plasticity_affine(self._h, self._n_plast - 1, affine, ac, aa, ab) != 0:
            raise ValueError("bad plasticity rule")

    def multipliers(self, proj):
        pr = self._projs[proj]
        nnz = pr.row_ptr[self._pops[pr.post].num]
        out = np.zeros(max(nnz, 1), dtype=np.float32)
        lib().orc_get_multipliers(self._h, proj, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        return out[:nnz]

    def spike_row(self, n):
        p = lib().orc_spike_row(self._h, n)
        return np.ctypeslib.as_array(p, shape=(self.total,)).copy()

    def steps_since_init(self):
        return lib().orc_steps_since_init(self._h)
