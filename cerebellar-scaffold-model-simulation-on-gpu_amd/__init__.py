"""Python binding of the MI355X cerebellar-scaffold engine (ctypes over the C ABI).

The classes mirror the reference simulator's host API
(SimulationModule/Scaffold/GPUScaffoldSimulation.h:41-46):

    env = NetworkEnvironment(num_of_dev=1, num_of_neuron_types=7,
                             num_of_connectivity_types=16, num_of_inputs=2)
    env.describe_synthetic(synth_preset(1))          # or set_neuron_params(...) etc.
    env.initialize(T_print=50, T_cpu_sim=25)          # InitializeNetworkEnv
    for t in range(0, 1000, 25):
        out = env.loop_n_steps(t, 25, trial=0)        # loop_n_steps
    env.reset(trial=0)                                # ResetNetworkEnv
    env.finalize()                                    # FinalizeNetworkEnv

Every call goes to the native library lib/libscaffold_mi355x.so; there is no
Python or CPU fallback, and a missing library or device raises ScaffoldError.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SCAFFOLD_LIB selects another build of the same library (e.g. the host-sanitizer
# build of `make asan`)
LIB_PATH = os.environ.get("SCAFFOLD_LIB") or os.path.join(_HERE, "lib", "libscaffold_mi355x.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "scaffold.h")

OK, ERR_INVALID_ARG, ERR_STATE, ERR_IO, ERR_FORMAT, ERR_UNSUPPORTED, ERR_HIP, ERR_NO_DEVICE, \
    ERR_WEIGHT_RANGE, ERR_CAPACITY = range(10)
NORMAL, INPUT, OUTPUT = 0, 1, 2
WEIGHTS_F32, WEIGHTS_BF16, WEIGHTS_FP8 = 0, 1, 2
# plasticity rules and coefficient slots (STDP_RULE, STDP_COEFFICIENTS: struct_enum_def.h:120-121)
STDP_HEBB, STDP_TEACHER = 0, 1
A0, A1_I, A1_J, A2_JI, A2_IJ = range(5)


class ScaffoldError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[scaffold error {code}] {msg}")
        self.code = code


class StimFunc(ctypes.Structure):
    _fields_ = [("poisson_hz", ctypes.c_float), ("poisson_on", ctypes.c_double), ("poisson_off", ctypes.c_double),
                ("burst_hz", ctypes.c_float), ("burst_on", ctypes.c_double), ("burst_off", ctypes.c_double),
                ("periodic_hz", ctypes.c_float), ("periodic_start", ctypes.c_float),
                ("periodic_on", ctypes.c_double), ("periodic_off", ctypes.c_double),
                ("periodic_off_inclusive", ctypes.c_int), ("poisson_hz_from_freq", ctypes.c_int),
                ("seed", ctypes.c_uint64),
                ("sine_max_hz", ctypes.c_float), ("sine_cycle_s", ctypes.c_float), ("sine_shift", ctypes.c_float),
                ("sine_on", ctypes.c_double), ("sine_off", ctypes.c_double)]


class StdpAffine(ctypes.Structure):
    """scaffold_stdp_affine: coefficient = (float)(c + a * (b - (double)w))."""
    _fields_ = [("c", ctypes.c_double), ("a", ctypes.c_double), ("b", ctypes.c_double)]


class SynthConfig(ctypes.Structure):
    _fields_ = [("scale", ctypes.c_double), ("counts", ctypes.c_int32 * 8), ("seed", ctypes.c_uint64),
                ("weight_jitter", ctypes.c_int32), ("mf_rate_hz", ctypes.c_float), ("burst_rate_hz", ctypes.c_float),
                ("burst_on_ms", ctypes.c_double), ("burst_off_ms", ctypes.c_double), ("burst_n_mf", ctypes.c_int32),
                ("burst_fraction", ctypes.c_double), ("stim_seed", ctypes.c_uint64), ("max_delay", ctypes.c_int32),
                ("reference_delays", ctypes.c_int32), ("connectome", ctypes.c_int32)]


class Info(ctypes.Structure):
    _fields_ = [("n_populations", ctypes.c_int32), ("n_projections", ctypes.c_int32), ("n_stimuli", ctypes.c_int32),
                ("total_neurons", ctypes.c_int32), ("device_neurons", ctypes.c_int32), ("T_print", ctypes.c_int32),
                ("T_cpu_sim", ctypes.c_int32), ("ring_rows", ctypes.c_int32), ("max_delay", ctypes.c_int32),
                ("total_synapses", ctypes.c_int64), ("steps_since_init", ctypes.c_int64),
                ("current_step", ctypes.c_int32), ("n_output_neurons", ctypes.c_int32),
                ("device_bytes", ctypes.c_int64), ("algorithmic_bytes_per_step", ctypes.c_int64),
                ("n_tasks", ctypes.c_int32), ("n_groups", ctypes.c_int32), ("synapses_16bit", ctypes.c_int64),
                ("n_blocks", ctypes.c_int32), ("n_staged_blocks", ctypes.c_int32),
                ("lane_synapses", ctypes.c_int32), ("stage_ratio", ctypes.c_int32),
                ("word_groups", ctypes.c_int64), ("word_group_synapses", ctypes.c_int64),
                ("stream_bytes_per_step", ctypes.c_int64)]


class PopulationInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("type", ctypes.c_int32), ("device_type", ctypes.c_int32),
                ("num", ctypes.c_int32), ("base_id", ctypes.c_int32), ("c_num", ctypes.c_int32),
                ("Cm", ctypes.c_float), ("El", ctypes.c_float), ("dt_ref", ctypes.c_float), ("Ie", ctypes.c_float),
                ("Vr", ctypes.c_float), ("Vth", ctypes.c_float), ("gL", ctypes.c_float),
                ("w", ctypes.c_float * 2), ("E", ctypes.c_float * 2), ("g_bar", ctypes.c_float * 2),
                ("tau", ctypes.c_float * 2), ("ref_step", ctypes.c_int32)]


class ProjectionInfo(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("pre", ctypes.c_int32), ("post", ctypes.c_int32),
                ("pre_num", ctypes.c_int32), ("post_num", ctypes.c_int32), ("delay", ctypes.c_int32),
                ("initial_weight", ctypes.c_float), ("w_bar", ctypes.c_float), ("nnz", ctypes.c_int64),
                ("max_conv", ctypes.c_int32), ("lane_mode", ctypes.c_int32), ("fp8_scale", ctypes.c_float),
                ("sum_scale", ctypes.c_float), ("plastic", ctypes.c_int32), ("grid_exp", ctypes.c_int32)]


class TuneResult(ctypes.Structure):
    _fields_ = [("lane_max_fanin", ctypes.c_int32), ("lane_synapses", ctypes.c_int32),
                ("graph_steps", ctypes.c_int32), ("n_tried", ctypes.c_int32), ("ms_per_step", ctypes.c_double),
                ("ms_default", ctypes.c_double), ("index_bits", ctypes.c_int32), ("stage_spikes", ctypes.c_int32),
                ("word_groups", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class StimInfo(ctypes.Structure):
    _fields_ = [("pop", ctypes.c_int32), ("n_entries", ctypes.c_int32), ("n_targets", ctypes.c_int32),
                ("func", StimFunc)]


_lib = None


def load_library(path=None):
    """Load the native engine; raises ScaffoldError when it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ScaffoldError(ERR_UNSUPPORTED, f"native library missing: {p} (run __graft_entry__.build())")
    lib = ctypes.CDLL(p)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    i32, i64, u64, f32, f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_float, ctypes.c_double
    cs = ctypes.c_char_p
    sig = {
        "scaffold_last_error": (cs, []),
        "scaffold_version": (cs, []),
        "scaffold_stim_func_reference": (ctypes.c_int, [ctypes.c_int, u64, P(StimFunc)]),
        "scaffold_create_network_env": (ctypes.c_int, [P(vp), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                       ctypes.c_int, ctypes.c_int]),
        "scaffold_set_neuron_params": (ctypes.c_int, [vp, ctypes.c_int, cs, ctypes.c_int, ctypes.c_int] + [f32] * 7
                                       + [ctypes.c_int, ctypes.c_int, P(f64), P(ctypes.c_int)]),
        "scaffold_set_connectivity_params": (ctypes.c_int, [vp, ctypes.c_int, cs] + [ctypes.c_int] * 4
                                             + [f32, f32, ctypes.c_int, P(ctypes.c_int)]),
        "scaffold_set_connectivity_csr": (ctypes.c_int, [vp, ctypes.c_int] + [ctypes.c_int] * 4
                                          + [f32, f32, P(i64), P(i32), P(f32), P(ctypes.c_int)]),
        "scaffold_set_plasticity": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(f32),
                                                   ctypes.c_uint32, ctypes.c_int, P(ctypes.c_int)]),
        "scaffold_set_plasticity_affine": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(StdpAffine),
                                                          ctypes.c_uint32, ctypes.c_int, P(ctypes.c_int)]),
        "scaffold_set_input_stim": (ctypes.c_int, [vp, ctypes.c_int, cs, P(StimFunc), P(ctypes.c_int)]),
        "scaffold_set_input_stim_ids": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_uint32), ctypes.c_int, P(i64),
                                                       P(StimFunc), P(ctypes.c_int)]),
        "scaffold_set_option": (ctypes.c_int, [vp, cs, i64]),
        "scaffold_get_option": (ctypes.c_int, [vp, cs, P(i64)]),
        "scaffold_initialize_network_env": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "scaffold_loop_n_steps": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_void_p)]),
        "scaffold_set_input_freq": (ctypes.c_int, [vp, ctypes.c_void_p]),  # float array (address)
        "scaffold_reset_network_env": (ctypes.c_int, [vp, ctypes.c_int]),
        "scaffold_finalize_network_env": (ctypes.c_int, [vp]),
        "scaffold_describe_reference": (ctypes.c_int, [vp, cs, u64]),
        "scaffold_synth_config_preset": (ctypes.c_int, [ctypes.c_int, P(SynthConfig)]),
        "scaffold_describe_synthetic": (ctypes.c_int, [vp, P(SynthConfig)]),
        "scaffold_write_dat": (ctypes.c_int, [vp, ctypes.c_int, cs]),
        "scaffold_get_info": (ctypes.c_int, [vp, P(Info)]),
        "scaffold_get_population": (ctypes.c_int, [vp, ctypes.c_int, P(PopulationInfo)]),
        "scaffold_get_positions": (ctypes.c_int, [vp, ctypes.c_int, vp, vp]),
        "scaffold_get_projection": (ctypes.c_int, [vp, ctypes.c_int, P(ProjectionInfo)]),
        "scaffold_get_projection_csr": (ctypes.c_int, [vp, ctypes.c_int, P(i64), P(i32), P(f32)]),
        "scaffold_get_stimulus": (ctypes.c_int, [vp, ctypes.c_int, P(StimInfo), P(i64), P(ctypes.c_uint32)]),
        "scaffold_get_state": (ctypes.c_int, [vp, P(f32), P(f32), P(i32)]),
        "scaffold_set_state": (ctypes.c_int, [vp, P(f32), P(f32), P(i32)]),
        "scaffold_get_spike_row": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_uint8)]),
        "scaffold_get_counts": (ctypes.c_int, [vp, ctypes.c_int, P(i32), P(u64)]),
        "scaffold_get_fired": (ctypes.c_int, [vp, ctypes.c_int, P(i32), i32, P(i32)]),
        "scaffold_get_fired_noalloc": (ctypes.c_int, [vp, ctypes.c_int, P(i32), i32, P(i32)]),
        "scaffold_get_totals": (ctypes.c_int, [vp, P(u64), P(u64)]),
        "scaffold_write_spike_text": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, cs]),
        "scaffold_write_output_spike_text": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, cs]),
        "scaffold_record_spikes": (ctypes.c_int, [vp, cs]),
        "scaffold_step_timer": (ctypes.c_int, [vp, ctypes.c_int, P(f64), P(i64)]),
        "scaffold_time_steps": (ctypes.c_int, [vp, ctypes.c_int, P(f64), P(f64)]),
        "scaffold_autotune": (ctypes.c_int, [vp, ctypes.c_int, vp]),
        "scaffold_time_layout": (ctypes.c_int, [vp, ctypes.c_int, P(f64), P(f64)]),
        "scaffold_batch_create": (ctypes.c_int, [P(vp)]),
        "scaffold_batch_add": (ctypes.c_int, [vp, vp]),
        "scaffold_batch_build": (ctypes.c_int, [vp]),
        "scaffold_batch_run": (ctypes.c_int, [vp, ctypes.c_int]),
        "scaffold_batch_time": (ctypes.c_int, [vp, ctypes.c_int, P(f64)]),
        "scaffold_batch_destroy": (ctypes.c_int, [vp]),
        "scaffold_copy_probe": (ctypes.c_int, [i64, ctypes.c_int, P(f64)]),
        "scaffold_encode_multipliers": (ctypes.c_int, [ctypes.c_int, P(f32), i64, P(f32), P(f32), P(f32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _check(rc):
    if rc != OK:
        raise ScaffoldError(rc, load_library().scaffold_last_error().decode(errors="replace"))


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def header_symbols(path=HEADER_PATH):
    """Names of every function the C header declares."""
    import re
    text = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(scaffold_\w+)\s*\(", text, re.M)))


def stim_func_reference(func_id, seed=0):
    f = StimFunc()
    _check(load_library().scaffold_stim_func_reference(func_id, seed, ctypes.byref(f)))
    return f


def synth_preset(which):
    c = SynthConfig()
    _check(load_library().scaffold_synth_config_preset(which, ctypes.byref(c)))
    return c


def encode_multipliers(precision, values):
    """(stored multipliers dequantized, FP8 scale, row-sum factor) of one
    projection's multipliers as the engine stores them with `precision`."""
    v = np.ascontiguousarray(values, dtype=np.float32)
    out = np.zeros_like(v)
    sc, ss = ctypes.c_float(), ctypes.c_float()
    _check(load_library().scaffold_encode_multipliers(precision, _ptr(v, ctypes.c_float), v.size,
                                                      _ptr(out, ctypes.c_float), ctypes.byref(sc), ctypes.byref(ss)))
    return out, sc.value, ss.value


def copy_probe(nbytes, iters=10):
    ms = ctypes.c_double()
    _check(load_library().scaffold_copy_probe(int(nbytes), int(iters), ctypes.byref(ms)))
    return ms.value


class NetworkEnvironment:
    """The reference's NetworkEnvironment with its lifecycle calls."""

    def __init__(self, num_of_dev=1, num_of_neuron_types=7, num_of_connectivity_types=16, num_of_inputs=2,
                 num_of_plasticities=0):
        self._lib = load_library()
        self._h = ctypes.c_void_p()
        self._freq_buf = None     # set_input_freq's rate array, one per stimulus (from initialize)
        self._train_view = None   # (address, array view) of the library's output-train buffer
        _check(self._lib.scaffold_create_network_env(ctypes.byref(self._h), num_of_dev, num_of_neuron_types,
                                                     num_of_connectivity_types, num_of_inputs,
                                                     num_of_plasticities))

    @property
    def handle(self):
        return self._h

    # ---------------------------------------------------------- description
    def set_neuron_params(self, type_, name, duplicate, num, Cm, El, dt_ref, Ie, Vr, Vth, gL, device_type,
                          channels=()):
        ch = np.asarray(channels, dtype=np.float64).reshape(-1)
        c_num = ch.size // 4
        idx = ctypes.c_int()
        _check(self._lib.scaffold_set_neuron_params(self._h, type_, name.encode(), duplicate, num, Cm, El, dt_ref,
                                                    Ie, Vr, Vth, gL, device_type, c_num,
                                                    _ptr(ch, ctypes.c_double) if c_num else None, ctypes.byref(idx)))
        return idx.value

    def set_connectivity_params(self, type_, filename, preNum, postNum, preType, postType, initial_weight, delay,
                                use_parallel_reduction=0):
        idx = ctypes.c_int()
        _check(self._lib.scaffold_set_connectivity_params(self._h, type_, str(filename).encode(), preNum, postNum,
                                                          preType, postType, initial_weight, delay,
                                                          use_parallel_reduction, ctypes.byref(idx)))
        return idx.value

    def set_connectivity_csr(self, type_, preNum, postNum, preType, postType, initial_weight, delay, row_ptr, col,
                             val=None):
        rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
        cl = np.ascontiguousarray(col, dtype=np.int32)
        vl = None if val is None else np.ascontiguousarray(val, dtype=np.float32)
        idx = ctypes.c_int()
        _check(self._lib.scaffold_set_connectivity_csr(self._h, type_, preNum, postNum, preType, postType,
                                                       initial_weight, delay, _ptr(rp, ctypes.c_int64),
                                                       _ptr(cl, ctypes.c_int32), _ptr(vl, ctypes.c_float),
                                                       ctypes.byref(idx)))
        return idx.value

    def set_plasticity(self, rule, target, teacher=-1, coefficients=None, time_window=50):
        """One plasticity rule (the table of Init_Plasticity, SynapticPlasticity.cu:20-33):
        `coefficients` maps slots (A0, A1_I, A1_J, A2_JI, A2_IJ) to constants; absent
        slots are the reference's NULL entries."""
        coefficients = coefficients or {}
        c = np.zeros(5, dtype=np.float32)
        present = 0
        for slot, v in coefficients.items():
            c[slot] = v
            present |= 1 << slot
        idx = ctypes.c_int()
        _check(self._lib.scaffold_set_plasticity(self._h, rule, target, teacher, _ptr(c, ctypes.c_float), present,
                                                 time_window, ctypes.byref(idx)))
        return idx.value

    def set_plasticity_affine(self, rule, target, teacher=-1, coefficients=None, time_window=50):
        """A plasticity rule with weight-dependent coefficients: `coefficients` maps
        slots to (c, a, b), the coefficient of a synapse being the float value of
        c + a * (b - (double)w) for its multiplier w before the update (the
        template's PF_PC_A1_J = 0.001*(1.0 - w) is (0, 0.001, 1.0))."""
        coefficients = coefficients or {}
        c = (StdpAffine * 5)()
        present = 0
        for slot, (c0, a, b) in coefficients.items():
            c[slot] = StdpAffine(c0, a, b)
            present |= 1 << slot
        idx = ctypes.c_int()
        _check(self._lib.scaffold_set_plasticity_affine(self._h, rule, target, teacher, c, present, time_window,
                                                        ctypes.byref(idx)))
        return idx.value

    def set_input_stim(self, neuron_index, id_file, func):
        idx = ctypes.c_int()
        _check(self._lib.scaffold_set_input_stim(self._h, neuron_index, str(id_file).encode(), ctypes.byref(func),
                                                 ctypes.byref(idx)))
        return idx.value

    def set_input_stim_ids(self, neuron_index, ids, func, fanout_ptr=None):
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        fp = None if fanout_ptr is None else np.ascontiguousarray(fanout_ptr, dtype=np.int64)
        n = ids.size if fp is None else fp.size - 1
        idx = ctypes.c_int()
        _check(self._lib.scaffold_set_input_stim_ids(self._h, neuron_index, _ptr(ids, ctypes.c_uint32), n,
                                                     _ptr(fp, ctypes.c_int64), ctypes.byref(func), ctypes.byref(idx)))
        return idx.value

    def describe_reference(self, data_dir, stim_seed=0):
        _check(self._lib.scaffold_describe_reference(self._h, str(data_dir).encode(), stim_seed))

    def describe_synthetic(self, cfg):
        _check(self._lib.scaffold_describe_synthetic(self._h, ctypes.byref(cfg)))

    def set_option(self, key, value):
        _check(self._lib.scaffold_set_option(self._h, key.encode(), int(value)))

    def get_option(self, key):
        v = ctypes.c_int64()
        _check(self._lib.scaffold_get_option(self._h, key.encode(), ctypes.byref(v)))
        return v.value

    def autotune(self, n_steps=200):
        """Time candidate layouts on private trial engines and set the fastest
        options (before initialize).  Returns the result as a dict."""
        r = TuneResult()
        _check(self._lib.scaffold_autotune(self._h, n_steps, ctypes.byref(r)))
        return {f: getattr(r, f) for f, _ in TuneResult._fields_}

    def time_layout(self, n_steps=100):
        """Time a private engine of the network under the current options (before
        initialize): (device ms per step, mean step-kernel ms)."""
        ms, km = ctypes.c_double(), ctypes.c_double()
        _check(self._lib.scaffold_time_layout(self._h, n_steps, ctypes.byref(ms), ctypes.byref(km)))
        return ms.value, km.value

    def write_dat(self, index, filename):
        _check(self._lib.scaffold_write_dat(self._h, index, str(filename).encode()))

    # ------------------------------------------------------------ lifecycle
    def initialize(self, T_print=50, T_cpu_sim=25):
        _check(self._lib.scaffold_initialize_network_env(self._h, T_print, T_cpu_sim))
        info = self.info()
        self._train_shape = (info.T_cpu_sim, info.n_output_neurons)  # fixed until finalize
        self._train_view = None
        if info.n_stimuli > 0:
            self._freq_buf = np.zeros(info.n_stimuli, dtype=np.float32)
            self._freq_ptr = self._freq_buf.ctypes.data

    def loop_n_steps(self, start_step, n_steps, trial=0, want_output=True):
        """Advance n_steps from start_step.  With want_output, returns the output
        spike train (T_cpu_sim x output neurons, uint8) of the last T_cpu_sim window
        the call completed, or None when it completed none."""
        out = ctypes.c_void_p()
        _check(self._lib.scaffold_loop_n_steps(self._h, start_step, n_steps, trial,
                                               ctypes.byref(out) if want_output else None))
        if not want_output or not out.value:
            return None
        T, n_out = self._train_shape
        if T * n_out == 0:
            return np.zeros((T, 0), dtype=np.uint8)
        if self._train_view is None or self._train_view[0] != out.value:
            buf = (ctypes.c_uint8 * (T * n_out)).from_address(out.value)
            self._train_view = (out.value, np.frombuffer(buf, dtype=np.uint8).reshape(T, n_out))
        return self._train_view[1].copy()

    def set_input_freq(self, freq_list):
        """SetInputFreq: one rate (Hz) per registered stimulus, in registration order."""
        buf = self._freq_buf
        if buf is None or len(freq_list) != buf.size:
            f = np.ascontiguousarray(freq_list, dtype=np.float32)
            _check(self._lib.scaffold_set_input_freq(self._h, f.ctypes.data))
            return
        buf[:] = freq_list
        _check(self._lib.scaffold_set_input_freq(self._h, self._freq_ptr))

    def reset(self, trial=0):
        _check(self._lib.scaffold_reset_network_env(self._h, trial))

    def finalize(self):
        if self._h:
            _check(self._lib.scaffold_finalize_network_env(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            if self._h:
                self._lib.scaffold_finalize_network_env(self._h)
        except Exception:
            pass

    # -------------------------------------------------------------- queries
    def info(self):
        i = Info()
        _check(self._lib.scaffold_get_info(self._h, ctypes.byref(i)))
        return i

    def population(self, index):
        p = PopulationInfo()
        _check(self._lib.scaffold_get_population(self._h, index, ctypes.byref(p)))
        return p

    def positions(self, index):
        """Cell positions (x, z) in um of a population of a synthetic network."""
        n = self.population(index).num
        x, z = np.zeros(n, np.float32), np.zeros(n, np.float32)
        _check(self._lib.scaffold_get_positions(self._h, index, x.ctypes.data, z.ctypes.data))
        return x, z

    def projection(self, index):
        p = ProjectionInfo()
        _check(self._lib.scaffold_get_projection(self._h, index, ctypes.byref(p)))
        return p

    def projection_csr(self, index):
        pi = self.projection(index)
        rp = np.zeros(pi.post_num + 1, dtype=np.int64)
        col = np.zeros(max(pi.nnz, 1), dtype=np.int32)
        val = np.zeros(max(pi.nnz, 1), dtype=np.float32)
        _check(self._lib.scaffold_get_projection_csr(self._h, index, _ptr(rp, ctypes.c_int64),
                                                     _ptr(col, ctypes.c_int32), _ptr(val, ctypes.c_float)))
        return rp, col[:pi.nnz], val[:pi.nnz]

    def stimulus(self, index):
        s = StimInfo()
        _check(self._lib.scaffold_get_stimulus(self._h, index, ctypes.byref(s), None, None))
        fp = np.zeros(s.n_entries + 1, dtype=np.int64)
        tg = np.zeros(max(s.n_targets, 1), dtype=np.uint32)
        _check(self._lib.scaffold_get_stimulus(self._h, index, ctypes.byref(s), _ptr(fp, ctypes.c_int64),
                                               _ptr(tg, ctypes.c_uint32)))
        return s, fp, tg[:s.n_targets]

    def state(self):
        n = self.info().total_neurons
        V = np.zeros(n, np.float32)
        g = np.zeros(2 * n, np.float32)
        r = np.zeros(n, np.int32)
        _check(self._lib.scaffold_get_state(self._h, _ptr(V, ctypes.c_float), _ptr(g, ctypes.c_float),
                                            _ptr(r, ctypes.c_int32)))
        return V, g, r

    def set_state(self, V=None, g=None, refr=None):
        V = None if V is None else np.ascontiguousarray(V, np.float32)
        g = None if g is None else np.ascontiguousarray(g, np.float32)
        refr = None if refr is None else np.ascontiguousarray(refr, np.int32)
        _check(self._lib.scaffold_set_state(self._h, _ptr(V, ctypes.c_float), _ptr(g, ctypes.c_float),
                                            _ptr(refr, ctypes.c_int32)))

    def spike_row(self, step):
        n = self.info().total_neurons
        s = np.zeros(n, np.uint8)
        _check(self._lib.scaffold_get_spike_row(self._h, step, _ptr(s, ctypes.c_uint8)))
        return s

    def counts(self, step):
        n = self.info().n_populations
        c = np.zeros(n, np.int32)
        ev = ctypes.c_uint64()
        _check(self._lib.scaffold_get_counts(self._h, step, _ptr(c, ctypes.c_int32), ctypes.byref(ev)))
        return c, ev.value

    def totals(self):
        """(spikes per population, synaptic events) accumulated since initialization."""
        n = self.info().n_populations
        sp = np.zeros(n, np.uint64)
        ev = ctypes.c_uint64()
        _check(self._lib.scaffold_get_totals(self._h, _ptr(sp, ctypes.c_uint64), ctypes.byref(ev)))
        return sp, ev.value

    def fired(self, step, cap=None):
        cap = self.info().total_neurons if cap is None else cap
        ids = np.zeros(max(cap, 1), np.int32)
        n = ctypes.c_int32()
        _check(self._lib.scaffold_get_fired(self._h, step, _ptr(ids, ctypes.c_int32), cap, ctypes.byref(n)))
        return ids[:min(n.value, cap)], n.value

    def fired_noalloc(self, step, cap=None):
        """fired() through the library's static scratch (no device allocation)."""
        cap = self.info().total_neurons if cap is None else cap
        ids = np.zeros(max(cap, 1), np.int32)
        n = ctypes.c_int32()
        _check(self._lib.scaffold_get_fired_noalloc(self._h, step, _ptr(ids, ctypes.c_int32), cap, ctypes.byref(n)))
        return ids[:min(n.value, cap)], n.value

    def write_spike_text(self, step0, n, filename):
        """The reference printer thread's lines (device populations) of n steps."""
        _check(self._lib.scaffold_write_spike_text(self._h, step0, n, str(filename).encode()))

    def write_output_spike_text(self, step0, n, filename):
        """The reference host thread's lines (OUTPUT populations) of n steps."""
        _check(self._lib.scaffold_write_output_spike_text(self._h, step0, n, str(filename).encode()))

    def record_spikes(self, directory):
        """The reference's PRINT mode: loop_n_steps writes spike_dev0_{0,1}.dat and
        spike_host{0,1}.dat into `directory` as it runs; reset() merges them into
        spike_dev0.dat and spike_result_trial<trial>.dat.  None stops recording."""
        _check(self._lib.scaffold_record_spikes(self._h, None if directory is None else str(directory).encode()))

    def step_timer(self, on=-1):
        """Step-kernel timer: on=n >= 1 clears and starts it, timing one run of
        steps in n; 0 stops it, -1 reads it.  Returns (milliseconds measured, step
        launches covered)."""
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        _check(self._lib.scaffold_step_timer(self._h, on, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def time_steps(self, n_steps, kernel_mean=False):
        total = ctypes.c_double()
        km = ctypes.c_double()
        _check(self._lib.scaffold_time_steps(self._h, n_steps, ctypes.byref(total),
                                             ctypes.byref(km) if kernel_mean else None))
        return (total.value, km.value) if kernel_mean else total.value


class Batch:
    """Initialized environments advanced together by one step kernel per step."""

    def __init__(self, envs):
        self._lib = load_library()
        self._h = ctypes.c_void_p()
        _check(self._lib.scaffold_batch_create(ctypes.byref(self._h)))
        self._envs = list(envs)
        for e in self._envs:
            _check(self._lib.scaffold_batch_add(self._h, e.handle))
        _check(self._lib.scaffold_batch_build(self._h))

    def run(self, n_steps):
        _check(self._lib.scaffold_batch_run(self._h, n_steps))

    def time(self, n_steps):
        ms = ctypes.c_double()
        _check(self._lib.scaffold_batch_time(self._h, n_steps, ctypes.byref(ms)))
        return ms.value

    def destroy(self):
        if self._h:
            _check(self._lib.scaffold_batch_destroy(self._h))
            self._h = ctypes.c_void_p()
