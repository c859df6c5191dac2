"""Python mirror of the CULS AIG manager surface over libgls.so (ctypes).

``Aig`` follows the reference's ``AIGMan`` (src/aig_manager.h:8-100): read /
write an AIGER file, ``strash``, ``rewrite(use_zeros)``, stats as printed by
``ps``.  All compute goes through the C ABI in include/gls.h; if the library is
missing this module raises instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgls.so")

GLS_LIT_AIGER = 0
GLS_LIT_CULS = 1
GLS_ST_CPU = 1
GLS_ST_WAVES = 2
GLS_ST_LEVELS = 4
GLS_RW_COOP_CUTS = 2
GLS_RW_DATAFLOW_CUTS = 4
GLS_RW_XCD_CUTS = 8
GLS_RW_LEVEL_CUTS = 16
GLS_RW_PAIR_CUTS = 32
GLS_RW_FULL_STRASH = 64
GLS_RW_GRAPH_CUTS = 128
GLS_RW_PLAN_CUTS = 256
GLS_RW_QUAD_CUTS = 512


class GlsError(RuntimeError):
    pass


class JobInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("nPIs", ctypes.c_int32), ("nPOs", ctypes.c_int32),
                ("nAnds", ctypes.c_int32), ("levels", ctypes.c_int32), ("nAnds_in", ctypes.c_int32),
                ("footprint_bytes", ctypes.c_int64), ("submit_ms", ctypes.c_double), ("admit_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(ctypes.Structure):
    _fields_ = [("nPIs", ctypes.c_int32), ("nPOs", ctypes.c_int32), ("nAnds", ctypes.c_int32),
                ("nObjs", ctypes.c_int32), ("levels", ctypes.c_int32), ("on_device", ctypes.c_int32)]


class OpInfo(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("reorder_ms", ctypes.c_double), ("cuts_ms", ctypes.c_double),
                ("table_ms", ctypes.c_double), ("eval_ms", ctypes.c_double), ("select_ms", ctypes.c_double),
                ("build_ms", ctypes.c_double), ("strash_ms", ctypes.c_double),
                ("candidates", ctypes.c_int64), ("selected", ctypes.c_int64), ("new_nodes", ctypes.c_int64),
                ("rounds", ctypes.c_int64), ("levels_enumerated", ctypes.c_int64), ("est_gain", ctypes.c_int64),
                ("kernel_sample_ms", ctypes.c_double), ("kernel_samples", ctypes.c_int64),
                ("kernel_sample_nodes", ctypes.c_int64), ("hash_probes", ctypes.c_int64),
                ("cut_schedule", ctypes.c_int32), ("final_strash", ctypes.c_int32),
                ("strash_rounds", ctypes.c_int64), ("cut_graph_ms", ctypes.c_double),
                ("cut_graph_nodes", ctypes.c_int64), ("strash_sweeps", ctypes.c_int32), ("pad0", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(ctypes.Structure):
    _fields_ = [("cut_wide", ctypes.c_int32), ("rw_sweep_width", ctypes.c_int32),
                ("strash_table_factor", ctypes.c_int32), ("strash_wave_batch", ctypes.c_int32),
                ("balance_narrow_max", ctypes.c_int32), ("cut_levels", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


GLS_TUNE_RW, GLS_TUNE_RWZ, GLS_TUNE_STRASH, GLS_TUNE_BALANCE = 0, 1, 2, 3
GLS_NET_CONST1, GLS_NET_CI, GLS_NET_AND, GLS_NET_CO = 0, 1, 2, 3


class Net(ctypes.Structure):
    _fields_ = [("nObjs", ctypes.c_int32), ("nCis", ctypes.c_int32), ("nCos", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("kind", ctypes.c_void_p), ("fanin0", ctypes.c_void_p),
                ("fanin1", ctypes.c_void_p), ("compl0", ctypes.c_void_p), ("compl1", ctypes.c_void_p),
                ("ci", ctypes.c_void_p), ("co", ctypes.c_void_p)]


class CecInfo(ctypes.Structure):
    _fields_ = [("equivalent", ctypes.c_int32), ("exhaustive", ctypes.c_int32), ("po", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("pattern", ctypes.c_int64), ("patterns", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


GLS_SIM_EXHAUSTIVE_PIS = 16

_lib = None

_P = ctypes.c_void_p
_SIGS = {
    "gls_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
    "gls_ctx_destroy": (None, [_P]),
    "gls_last_error": (ctypes.c_char_p, [_P]),
    "gls_ctx_has_device": (ctypes.c_int, [_P]),
    "gls_aig_from_arrays": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P, _P, _P,
                                           ctypes.c_int32, ctypes.POINTER(_P)]),
    "gls_aig_to_arrays": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_int32]),
    "gls_aig_stats": (ctypes.c_int, [_P, ctypes.POINTER(Stats)]),
    "gls_aig_copy": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "gls_aig_free": (None, [_P]),
    "gls_aig_read_aiger": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.POINTER(_P)]),
    "gls_aig_read_aiger_mem": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int64, ctypes.POINTER(_P)]),
    "gls_aig_write_aiger": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "gls_strash": (ctypes.c_int, [_P, ctypes.c_int32]),
    "gls_rewrite": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "gls_aig_last_op": (ctypes.c_int, [_P, ctypes.POINTER(OpInfo)]),
    "gls_balance": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "gls_server_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(_P)]),
    "gls_server_submit": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P, _P, _P,
                                         ctypes.c_int32, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]),
    "gls_server_wait": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(JobInfo)]),
    "gls_server_result": (ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, ctypes.c_int32]),
    "gls_server_release": (ctypes.c_int, [_P, ctypes.c_int64]),
    "gls_server_last_error": (ctypes.c_char_p, [_P]),
    "gls_server_destroy": (None, [_P]),
    "gls_batch_strash": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.c_int32, ctypes.c_int32]),
    "gls_batch_rewrite": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "gls_batch_balance": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "gls_aig_from_net": (ctypes.c_int, [_P, ctypes.c_void_p, ctypes.POINTER(_P)]),
    "gls_ctx_get_tuning": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_void_p]),
    "gls_ctx_set_tuning": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_void_p]),
    "gls_autotune": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)]),
    "gls_reorder": (ctypes.c_int, [_P, _P]),
    "gls_enumerate_cuts": (ctypes.c_int, [_P, _P]),
    "gls_evaluate": (ctypes.c_int, [_P, ctypes.c_int32, _P]),
    "gls_npn_tables": (ctypes.c_int, [_P, _P, _P, _P]),
    "gls_cec": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_uint64, ctypes.POINTER(CecInfo)]),
    "gls_simulate": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32, _P]),
    "gls_gen_random": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_uint64, ctypes.POINTER(_P)]),
    "gls_gen_geometric": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(_P)]),
    "gls_gen_multiplier": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_double, ctypes.c_uint64,
                                          ctypes.POINTER(_P)]),
}


def lib():
    """Loads libgls.so (raises GlsError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GlsError("libgls.so is not built: run `python -m culs_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Context:
    """One HIP device + stream (device=-1: host-only: I/O and generators)."""

    def __init__(self, device=0):
        self._h = _P()
        rc = lib().gls_ctx_create(device, ctypes.byref(self._h))
        if rc != 0:
            raise GlsError("gls_ctx_create(%d) failed with %d" % (device, rc))
        self.device = device

    def check(self, rc, what):
        if rc != 0:
            msg = lib().gls_last_error(self._h)
            raise GlsError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    @property
    def has_device(self):
        return bool(lib().gls_ctx_has_device(self._h))

    def close(self):
        if self._h:
            lib().gls_ctx_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # constructors
    def from_arrays(self, nPIs, f0, f1, pos, convention=GLS_LIT_AIGER):
        f0 = np.ascontiguousarray(f0, dtype=np.uint32)
        f1 = np.ascontiguousarray(f1, dtype=np.uint32)
        pos = np.ascontiguousarray(pos, dtype=np.uint32)
        n = len(f0)
        h = _P()
        self.check(lib().gls_aig_from_arrays(self._h, nPIs, len(pos), n - nPIs - 1, _ptr(f0), _ptr(f1),
                                             _ptr(pos) if len(pos) else None, convention, ctypes.byref(h)),
                   "gls_aig_from_arrays")
        return Aig(self, h)

    def read_aiger(self, path):
        h = _P()
        self.check(lib().gls_aig_read_aiger(self._h, path.encode(), ctypes.byref(h)), "gls_aig_read_aiger")
        return Aig(self, h)

    def read_aiger_bytes(self, data):
        h = _P()
        self.check(lib().gls_aig_read_aiger_mem(self._h, data, len(data), ctypes.byref(h)), "read_aiger_mem")
        return Aig(self, h)

    def gen_random(self, nPIs, nAnds, nPOs, window, seed):
        h = _P()
        self.check(lib().gls_gen_random(self._h, nPIs, nAnds, nPOs, window, seed, ctypes.byref(h)), "gen_random")
        return Aig(self, h)

    def gen_geometric(self, nPIs, nAnds, nPOs, window, p, seed):
        h = _P()
        self.check(lib().gls_gen_geometric(self._h, nPIs, nAnds, nPOs, window, p, seed, ctypes.byref(h)),
                   "gen_geometric")
        return Aig(self, h)

    def _handles(self, aigs):
        arr = (_P * max(1, len(aigs)))()
        for i, a in enumerate(aigs):
            arr[i] = a._h
        return arr

    def batch_strash(self, aigs):
        """Strash every AIG of the list in one device pass (results as if alone)."""
        self.check(lib().gls_batch_strash(self._h, self._handles(aigs), len(aigs), 0), "gls_batch_strash")
        return aigs

    def batch_rewrite(self, aigs, use_zeros=False, cut_mode=0):
        """One rw pass over every AIG of the list, packed into one arena."""
        self.check(lib().gls_batch_rewrite(self._h, self._handles(aigs), len(aigs), 1 if use_zeros else 0, cut_mode),
                   "gls_batch_rewrite")
        return aigs

    def batch_balance(self, aigs, update_level=True, shuffle=False):
        """`b` over every AIG of the list in one device pass (results as if
        alone); shuffle (`b -s`) is refused by the batched pass."""
        self.check(lib().gls_batch_balance(self._h, self._handles(aigs), len(aigs), 0 if shuffle else 1,
                                           1 if update_level else 0), "gls_batch_balance")
        return aigs

    def from_net(self, kind, fanin0, fanin1, compl0, compl1, ci, co):
        """gls_aig_from_net: an object network (GLS_NET_* kinds, fanin object
        indices and complement bits, CI and CO object lists) numbered as the
        ABC patch's gget numbers it."""
        kind = np.ascontiguousarray(kind, dtype=np.uint8)
        f0 = np.ascontiguousarray(fanin0, dtype=np.int32)
        f1 = np.ascontiguousarray(fanin1, dtype=np.int32)
        c0 = np.ascontiguousarray(compl0, dtype=np.uint8)
        c1 = np.ascontiguousarray(compl1, dtype=np.uint8)
        ci = np.ascontiguousarray(ci, dtype=np.int32)
        co = np.ascontiguousarray(co, dtype=np.int32)
        net = Net(len(kind), len(ci), len(co), 0, _ptr(kind), _ptr(f0), _ptr(f1), _ptr(c0), _ptr(c1),
                  _ptr(ci) if len(ci) else None, _ptr(co) if len(co) else None)
        h = _P()
        self.check(lib().gls_aig_from_net(self._h, ctypes.byref(net), ctypes.byref(h)), "gls_aig_from_net")
        return Aig(self, h)

    def get_tuning(self, objects):
        """The pass settings used for graphs (or batch arenas) of `objects` objects."""
        t = Tuning()
        self.check(lib().gls_ctx_get_tuning(self._h, objects, ctypes.byref(t)), "gls_ctx_get_tuning")
        return t.as_dict()

    def set_tuning(self, objects, **settings):
        """Changes some settings for that size bucket (objects < 0: every bucket)."""
        t = Tuning(**self.get_tuning(max(objects, 1)))
        for k, v in settings.items():
            setattr(t, k, v)
        self.check(lib().gls_ctx_set_tuning(self._h, objects, ctypes.byref(t)), "gls_ctx_set_tuning")

    def autotune(self, aigs, command=GLS_TUNE_RW, reps=2):
        """Times `command` on device copies of the AIGs under candidate settings
        and keeps the fastest for their size bucket -> (settings, best_ms, start_ms)."""
        t = Tuning()
        best, start = ctypes.c_double(0), ctypes.c_double(0)
        self.check(lib().gls_autotune(self._h, self._handles(aigs), len(aigs), command, reps, ctypes.byref(t),
                                      ctypes.byref(best), ctypes.byref(start)), "gls_autotune")
        return t.as_dict(), best.value, start.value

    def gen_multiplier(self, nbits, redundant, seed):
        h = _P()
        self.check(lib().gls_gen_multiplier(self._h, nbits, redundant, seed, ctypes.byref(h)), "gen_multiplier")
        return Aig(self, h)


class Aig:
    def __init__(self, ctx, handle):
        self.ctx = ctx
        self._h = handle

    def free(self):
        if self._h:
            lib().gls_aig_free(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def stats(self):
        s = Stats()
        self.ctx.check(lib().gls_aig_stats(self._h, ctypes.byref(s)), "gls_aig_stats")
        return s

    @property
    def n_ands(self):
        return self.stats().nAnds

    def arrays(self, convention=GLS_LIT_AIGER):
        """(nPIs, fanin0, fanin1, pos) as uint32 numpy arrays indexed by id."""
        s = self.stats()
        f0 = np.zeros(s.nObjs, np.uint32)
        f1 = np.zeros(s.nObjs, np.uint32)
        pos = np.zeros(max(s.nPOs, 1), np.uint32)
        self.ctx.check(lib().gls_aig_to_arrays(self._h, _ptr(f0), _ptr(f1), _ptr(pos), convention),
                       "gls_aig_to_arrays")
        return s.nPIs, f0, f1, pos[:s.nPOs]

    def copy(self):
        h = _P()
        self.ctx.check(lib().gls_aig_copy(self._h, ctypes.byref(h)), "gls_aig_copy")
        return Aig(self.ctx, h)

    def write_aiger(self, path):
        self.ctx.check(lib().gls_aig_write_aiger(self._h, path.encode()), "gls_aig_write_aiger")

    def strash(self, cpu=False, schedule=0):
        """Device strash; cpu=True is `st -c` (host strash without trivial
        folding); schedule GLS_ST_WAVES / GLS_ST_LEVELS forces one of the two
        device schedules (same result)."""
        self.ctx.check(lib().gls_strash(self._h, GLS_ST_CPU if cpu else int(schedule)), "gls_strash")
        return self

    def rewrite(self, use_zeros=False, cut_mode=0, full_strash=False):
        """One rw pass (rw -z / rwz with use_zeros).  cut_mode schedules cut
        enumeration: 0 one launch per three levels, GLS_RW_PAIR_CUTS per two,
        GLS_RW_LEVEL_CUTS per level, GLS_RW_COOP_CUTS one cooperative launch,
        GLS_RW_DATAFLOW_CUTS fanin-driven, GLS_RW_XCD_CUTS one persistent
        kernel on one XCD; GLS_RW_GRAPH_CUTS or-ed into the first three
        captures the pass's cut launches in one HIP graph; full_strash
        finishes with the wave-sweep strash instead of the incremental one.
        The results are identical."""
        flags = cut_mode | (GLS_RW_FULL_STRASH if full_strash else 0)
        self.ctx.check(lib().gls_rewrite(self._h, 1 if use_zeros else 0, flags), "gls_rewrite")
        return self

    def balance(self, update_level=True, shuffle=False):
        """`b` (update_level: strash afterwards, as the command does; False is
        `gb`); shuffle: `b -s` (equal levels in thrust's shuffle order)."""
        self.ctx.check(lib().gls_balance(self._h, 0 if shuffle else 1, 1 if update_level else 0), "gls_balance")
        return self

    def last_op(self):
        info = OpInfo()
        self.ctx.check(lib().gls_aig_last_op(self._h, ctypes.byref(info)), "gls_aig_last_op")
        return info.as_dict()

    def reorder(self):
        s = self.stats()
        lev = np.zeros(max(s.nObjs, 1), np.uint32)
        self.ctx.check(lib().gls_reorder(self._h, _ptr(lev)), "gls_reorder")
        return lev[:self.stats().nObjs]

    def enumerate_cuts(self, cut_mode=0):
        s = self.stats()
        cuts = np.zeros(s.nObjs * 40, np.uint32)
        self.ctx.check(lib().gls_enumerate_cuts_sched(self._h, int(cut_mode), _ptr(cuts)), "gls_enumerate_cuts_sched")
        return cuts

    def evaluate(self, use_zeros=False):
        s = self.stats()
        best = np.zeros(s.nObjs * 4, np.int32)
        self.ctx.check(lib().gls_evaluate(self._h, 1 if use_zeros else 0, _ptr(best)), "gls_evaluate")
        return best

    def cec(self, other, words=64, seed=1):
        """Equivalence with `other` by device simulation (exhaustive for <= 16 PIs)."""
        info = CecInfo()
        self.ctx.check(lib().gls_cec(self._h, other._h, int(words), int(seed), ctypes.byref(info)), "gls_cec")
        return info.as_dict()

    def simulate(self, words=1, seed=1, exhaustive=False):
        """PO values over words*64 patterns: uint64 array [nPOs, words]."""
        s = self.stats()
        out = np.zeros((max(s.nPOs, 1), words), np.uint64)
        self.ctx.check(lib().gls_simulate(self._h, int(words), int(seed), 1 if exhaustive else 0, _ptr(out)),
                       "gls_simulate")
        return out[:s.nPOs]

    def ps(self):
        """The reference's `ps` line (src/aig_manager.cu:604-612)."""
        s = self.stats()
        return "AIG stats: i/o = %d/%d and = %d level = %d" % (s.nPIs, s.nPOs, s.nAnds, s.levels)


def npn_tables():
    ph = np.zeros(65536, np.uint8)
    pe = np.zeros(65536, np.uint8)
    cl = np.zeros(65536, np.uint8)
    p4 = np.zeros(96, np.uint8)
    rc = lib().gls_npn_tables(_ptr(ph), _ptr(pe), _ptr(cl), _ptr(p4))
    if rc != 0:
        raise GlsError("gls_npn_tables failed")
    return ph, pe, cl, p4.reshape(24, 4)


class Server:
    """Continuous-batching rewrite server (gls_server_*): submit(arrays, script)
    returns a job id; wait(job) -> info dict; result(job) -> arrays."""

    def __init__(self, device=0, hbm_budget=64 << 30, max_active=4096):
        self._h = _P()
        rc = lib().gls_server_create(device, hbm_budget, max_active, ctypes.byref(self._h))
        if rc != 0:
            raise GlsError("gls_server_create failed (%d)" % rc)

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().gls_server_last_error(self._h)
            raise GlsError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def submit(self, g, script):
        nPIs, f0, f1, pos = g
        f0 = np.ascontiguousarray(f0, dtype=np.uint32)
        f1 = np.ascontiguousarray(f1, dtype=np.uint32)
        pos = np.ascontiguousarray(pos, dtype=np.uint32)
        job = ctypes.c_int64(0)
        self._check(lib().gls_server_submit(self._h, nPIs, len(pos), len(f0) - nPIs - 1, _ptr(f0), _ptr(f1),
                                            _ptr(pos) if len(pos) else None, GLS_LIT_AIGER, script.encode(),
                                            ctypes.byref(job)), "gls_server_submit")
        return job.value

    def wait(self, job, timeout_ms=-1.0):
        info = JobInfo()
        self._check(lib().gls_server_wait(self._h, job, timeout_ms, ctypes.byref(info)), "gls_server_wait")
        return info.as_dict()

    def result(self, job):
        info = self.wait(job)
        n = info["nPIs"] + 1 + info["nAnds"]
        f0 = np.zeros(n, np.uint32)
        f1 = np.zeros(n, np.uint32)
        pos = np.zeros(max(1, info["nPOs"]), np.uint32)
        self._check(lib().gls_server_result(self._h, job, _ptr(f0), _ptr(f1), _ptr(pos), GLS_LIT_AIGER),
                    "gls_server_result")
        return info["nPIs"], f0, f1, pos[:info["nPOs"]]

    def release(self, job):
        self._check(lib().gls_server_release(self._h, job), "gls_server_release")

    def close(self):
        if self._h:
            lib().gls_server_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
