"""ctypes binding of the C ABI in include/mytrim_gpu.h (libmytrim_gpu.so).

The library must have been built (``python -c "import __graft_entry__ as g;
g.build()"`` or ``make -C mytrim_amd``); nothing here falls back to a CPU
path: every call goes to the HIP kernels and raises on failure.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Optional

import numpy as np

from .problem import Problem

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmytrim_gpu.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
lp = C.POINTER(C.c_longlong)

MT_FP64, MT_FP32, MT_FP32_SURROGATE = 0, 1, 2
MT_LOG_DEFECTS, MT_LOG_HISTORY, MT_LOG_PHONONS, MT_LOG_VACMAP, MT_LOG_TRANSPORT = 1, 2, 4, 8, 16
# one event record (mt_record); kinds: 1 V, 2 I, 3 S, 4 R (TrimDefectLog),
# 5 TrimHistory, 6 / 7 / 8 TrimPhononOut (spawn, dissipation, ion at rest)
RECORD = np.dtype([("pos", "<f8", 3), ("E", "<f8"), ("value", "<f8"), ("m", "<f8"), ("id", "<u8"),
                   ("hist", "<u4"), ("Z", "<i4"), ("gen", "<i4"), ("tag", "<i4"), ("kind", "<i4"),
                   ("pad", "<i4")])

# every symbol include/mytrim_gpu.h declares
EXPORTS = ["mt_last_error", "mt_version", "mt_ctx_create", "mt_ctx_destroy", "mt_ctx_set_tables",
           "mt_default_tables",
           "mt_ctx_set_target", "mt_ctx_set_tuning", "mt_run_cascades", "mt_run_cascades_log", "mt_autotune",
           "mt_trim_flight",
           "mt_run_cascades_noalloc",
           "mt_stopping_batch", "mt_scatter_batch", "mt_stopping_batch_arith", "mt_scatter_batch_arith",
           "mt_surrogate_norm", "mt_surrogate_train", "mt_surrogate_install",
           "mt_surrogate_eval", "mt_surrogate_grad", "mt_surrogate_ablate"]


class Target(C.Structure):
    _fields_ = [("n_mat", C.c_int), ("mat_rho", dp), ("mat_el_off", ip), ("el_Z", ip), ("el_m", dp),
                ("el_t", dp), ("el_Edisp", dp), ("el_Elbind", dp), ("geo", C.c_int),
                ("w", C.c_double * 3), ("bc", C.c_int * 3), ("n_layers", C.c_int),
                ("layer_thick", dp), ("kn", C.c_int * 3), ("n_cl", C.c_int), ("cl_x", dp),
                ("cl_y", dp), ("cl_z", dp), ("cl_r", dp), ("tmin", C.c_double), ("tau", C.c_double),
                ("cw", C.c_double), ("scale", C.c_double), ("potential", C.c_int)]


class Primaries(C.Structure):
    _fields_ = [("n_hist", C.c_longlong), ("hist0", C.c_longlong), ("pka_uniform", C.c_int),
                ("Z", ip), ("m", dp), ("E", dp), ("pos", dp), ("dir", dp), ("Ef", C.c_double),
                ("seed", C.c_uint32)]


class RunOpts(C.Structure):
    _fields_ = [("follow_max_gen", C.c_int), ("kp", C.c_int), ("arith", C.c_int),
                ("range_x0", C.c_double), ("range_bw", C.c_double), ("range_nb", C.c_int),
                ("lat_y0", C.c_double), ("lat_z0", C.c_double), ("lat_bw", C.c_double),
                ("lat_nb", C.c_int), ("vac_x0", C.c_double), ("vac_bw", C.c_double),
                ("vac_nb", C.c_int), ("per_history", C.c_int)]


COUNTERS = ["n_coll", "n_flight", "vac", "repl", "subst", "inter", "lost", "vacuum",
            "rutherford", "newton", "disp"]
ENERGIES = ["E0", "Eel", "Enuc", "Eunf", "Elost", "Edrop", "Evac"]


class LogSpec(C.Structure):
    _fields_ = [("mask", C.c_int), ("records", C.c_void_p), ("max_records", C.c_longlong),
                ("n_records", C.c_longlong), ("vacmap_Z", C.c_int * 3), ("vacmap", ip),
                ("transport", lp)]


class Tallies(C.Structure):
    _fields_ = ([(k, C.c_longlong) for k in COUNTERS] + [(k, C.c_double) for k in ENERGIES] + [
        ("range_n", C.c_longlong), ("range_sum", C.c_double), ("range_sum2", C.c_double),
        ("range_hist", lp), ("lat_y_hist", lp), ("lat_z_hist", lp), ("vac_hist", lp),
        ("repl_hist", lp), ("h_vac", ip), ("h_coll", lp), ("h_flights", ip), ("h_x", dp),
        ("h_y", dp), ("h_z", dp), ("h_state", ip), ("h_Eel", dp), ("kernel_ms", C.c_double),
        ("grid", C.c_int), ("block", C.c_int), ("waves", C.c_int), ("queue_peak", C.c_longlong),
        ("queue_ops", C.c_longlong), ("surrogate", C.c_longlong), ("coll_primary", C.c_longlong),
        ("sclk_mhz", C.c_double), ("half_ms", C.c_double), ("flights_at_half", C.c_longlong),
        ("admit_ms", C.c_double), ("flights_at_admit", C.c_longlong)])


class Tuning(C.Structure):
    _fields_ = [("block", C.c_int), ("blocks_per_cu", C.c_int), ("stack_cap", C.c_int),
                ("queue_cap", C.c_longlong), ("lds_hist", C.c_int), ("graph", C.c_int),
                ("primary_waves64", C.c_int)]


class SurrogateOpts(C.Structure):
    _fields_ = [("n_ions", C.c_int), ("ion_Z", C.POINTER(C.c_int)), ("ion_m", C.POINTER(C.c_double)),
                ("E_min", C.c_double), ("E_max", C.c_double), ("eps_max", C.c_double),
                ("n_train", C.c_longlong), ("epochs", C.c_int), ("batch", C.c_int), ("lr", C.c_double),
                ("seed", C.c_uint32), ("bf16", C.c_int), ("inputs", C.c_int), ("sampling", C.c_int),
                ("eps_min", C.c_double)]


N_SURROGATE_PARAMS = 4481
SURROGATE_NORM_KEYS = ["e_mid", "e_inv_half", "lb_mid", "lb_inv_half", "b_mid", "b_inv_half", "b_floor",
                       "y_off", "y_scale", "eps_lo", "eps_hi", "b_hi"]


class Ion(C.Structure):
    _fields_ = [("Z", C.c_int), ("m", C.c_double), ("E", C.c_double), ("pos", C.c_double * 3),
                ("dir", C.c_double * 3), ("gen", C.c_int), ("tag", C.c_int), ("state", C.c_int),
                ("Ef", C.c_double)]


class Event(C.Structure):
    _fields_ = [("pka_E", C.c_double), ("pka_pos", C.c_double * 3), ("pka_dir", C.c_double * 3),
                ("ls", C.c_double), ("dee", C.c_double), ("den", C.c_double), ("rec_E", C.c_double),
                ("rec_pos", C.c_double * 3), ("rec_dir", C.c_double * 3), ("material", C.c_int),
                ("element", C.c_int), ("pka_state", C.c_int), ("fate", C.c_int),
                ("vacancy", C.c_int), ("mat_tag", C.c_int)]


_lib = None


def lib():
    """Load libmytrim_gpu.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmytrim_gpu.so not built: run __graft_entry__.build() or "
                               "make -C mytrim_amd")
        L = C.CDLL(LIB_PATH)
        L.mt_last_error.restype = C.c_char_p
        L.mt_version.restype = C.c_char_p
        L.mt_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        L.mt_ctx_destroy.argtypes = [C.c_void_p]
        L.mt_ctx_set_tables.argtypes = [C.c_void_p, dp, C.c_int, C.c_int]
        L.mt_default_tables.argtypes = [dp, C.c_int, C.c_int]
        L.mt_ctx_set_target.argtypes = [C.c_void_p, C.POINTER(Target)]
        L.mt_ctx_set_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
        L.mt_run_cascades_log.argtypes = [C.c_void_p, C.POINTER(Primaries), C.POINTER(RunOpts),
                                          C.POINTER(LogSpec), C.POINTER(Tallies)]
        L.mt_run_cascades.argtypes = [C.c_void_p, C.POINTER(Primaries), C.POINTER(RunOpts),
                                      C.POINTER(Tallies)]
        L.mt_run_cascades_noalloc.argtypes = [C.c_void_p, C.POINTER(Primaries), C.POINTER(RunOpts),
                                              C.POINTER(Tallies)]
        L.mt_autotune.argtypes = [C.c_void_p, C.POINTER(Primaries), C.POINTER(RunOpts), C.c_longlong,
                                  C.c_int, C.POINTER(Tuning), dp]
        L.mt_trim_flight.argtypes = [C.c_void_p, C.POINTER(Ion), dp, C.c_longlong, lp,
                                     C.POINTER(Event), C.c_int, ip]
        L.mt_stopping_batch.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, dp,
                                        C.c_longlong, dp]
        L.mt_scatter_batch.argtypes = [C.c_void_p, C.c_int, dp, dp, C.c_longlong, dp, dp, dp, ip]
        L.mt_stopping_batch_arith.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, dp,
                                              C.c_longlong, dp]
        L.mt_scatter_batch_arith.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, dp, C.c_longlong, dp, dp,
                                             dp, ip]
        fp = C.POINTER(C.c_float)
        so = C.POINTER(SurrogateOpts)
        L.mt_surrogate_train.argtypes = [C.c_void_p, so, fp, dp]
        L.mt_surrogate_norm.argtypes = [C.c_void_p, so, fp]
        L.mt_surrogate_install.argtypes = [C.c_void_p, so, fp]
        L.mt_surrogate_eval.argtypes = [C.c_void_p, dp, dp, C.c_longlong, dp]
        L.mt_surrogate_grad.argtypes = [C.c_void_p, so, fp, dp, dp, dp, C.c_int, fp, dp]
        L.mt_surrogate_ablate.argtypes = [C.c_void_p, C.c_int, dp, dp, C.c_longlong, C.c_int, dp, dp]
        _lib = L
    return _lib


class MtError(RuntimeError):
    pass


def default_tables() -> np.ndarray:
    """The SCOEF/SLFCTR rows built into the library (92 x 16)."""
    rows = np.zeros((92, 16))
    _check(lib().mt_default_tables(_p(rows, dp), 92, 16), "mt_default_tables")
    return rows


def _check(rc: int, what: str):
    if rc != 0:
        raise MtError("%s failed (%d): %s" % (what, rc, lib().mt_last_error().decode()))


def _p(a, t):
    return a.ctypes.data_as(t)


class Engine:
    """One device context (mt_ctx): target tables on the GPU plus its stream."""

    def __init__(self, device: int = 0):
        self._keep = []
        h = C.c_void_p()
        _check(lib().mt_ctx_create(device, C.byref(h)), "mt_ctx_create")
        self.h = h
        self.problem: Optional[Problem] = None
        self._target_key = None

    def close(self):
        if self.h:
            lib().mt_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, block=0, blocks_per_cu=0, stack_cap=0, queue_cap=0, lds_hist=0, graph=0, primary_waves64=0):
        t = Tuning(block, blocks_per_cu, stack_cap, queue_cap, lds_hist, graph, primary_waves64)
        _check(lib().mt_ctx_set_tuning(self.h, C.byref(t)), "mt_ctx_set_tuning")

    def set_tables(self, rows: np.ndarray):
        """Replace the SCOEF/SLFCTR rows (92 x 16, the layout of
        default_tables()); runs after this use them."""
        rows = np.ascontiguousarray(rows, np.float64)
        _check(lib().mt_ctx_set_tables(self.h, _p(rows, dp), rows.shape[0], rows.shape[1]),
               "mt_ctx_set_tables")

    def set_target(self, prob: Problem):
        rho, off, Z, m, t, ed, eb = prob.flat_materials()
        g = prob.geometry
        layers = np.ascontiguousarray(g.layers, np.float64) if len(g.layers) else np.zeros(1)
        cl = np.ascontiguousarray(g.clusters if g.clusters is not None else np.zeros((1, 4)))
        cols = [np.ascontiguousarray(cl[:, i]) for i in range(4)]
        keep = [rho, off, Z, m, t, ed, eb, layers] + cols
        tg = Target()
        tg.n_mat = len(rho)
        tg.mat_rho, tg.mat_el_off, tg.el_Z = _p(rho, dp), _p(off, ip), _p(Z, ip)
        tg.el_m, tg.el_t, tg.el_Edisp, tg.el_Elbind = _p(m, dp), _p(t, dp), _p(ed, dp), _p(eb, dp)
        tg.geo = g.kind
        tg.w[:] = [float(x) for x in g.w]
        tg.bc[:] = [int(x) for x in g.bc]
        tg.n_layers = len(g.layers)
        tg.layer_thick = _p(layers, dp)
        tg.kn[:] = [int(x) for x in g.kn]
        tg.n_cl = 0 if g.clusters is None else len(g.clusters)
        tg.cl_x, tg.cl_y, tg.cl_z, tg.cl_r = (_p(c, dp) for c in cols)
        tg.tmin, tg.tau, tg.cw, tg.scale = prob.tmin, prob.tau, prob.cw, prob.scale
        tg.potential = prob.potential
        _check(lib().mt_ctx_set_target(self.h, C.byref(tg)), "mt_ctx_set_target")
        del keep
        self.problem = prob
        self._target_key = self.target_key(prob)

    @staticmethod
    def target_key(prob: Problem):
        """What mt_ctx_set_target receives from a problem: materials, geometry
        and the collision knobs.  Batches of one problem (Problem.batch) share
        it; their primaries and history range go with each run instead.  The
        bubble array is keyed by identity (the engine keeps the last problem,
        and with it the array, alive)."""
        g = prob.geometry
        mats = tuple((m.rho, tuple((e.Z, e.m, e.t, e.Edisp, e.Elbind) for e in m.elements)) for m in prob.materials)
        cl = None if g.clusters is None else (id(g.clusters), g.clusters.shape)
        return (mats, g.kind, tuple(float(x) for x in g.w), tuple(int(x) for x in g.bc),
                tuple(float(x) for x in g.layers), tuple(int(x) for x in g.kn), cl,
                prob.tmin, prob.tau, prob.cw, prob.scale, prob.potential)

    def _launch_args(self, prob: Problem, per_history: bool, arith: int):
        if self.problem is not prob and self._target_key != self.target_key(prob):
            self.set_target(prob)
        p = prob.primaries
        keep = [np.ascontiguousarray(p.Z, np.int32), np.ascontiguousarray(p.m, np.float64),
                np.ascontiguousarray(p.E, np.float64), np.ascontiguousarray(p.pos, np.float64),
                np.ascontiguousarray(p.dir, np.float64)]
        pZ, pm, pE, ppos, pdir = keep
        pk = Primaries(prob.n_hist, prob.hist0, 1 if p.uniform else 0, _p(pZ, ip), _p(pm, dp),
                       _p(pE, dp), _p(ppos, dp), _p(pdir, dp), prob.Ef, prob.seed & 0xFFFFFFFF)
        o = RunOpts(prob.follow_max_gen, int(prob.kp), arith, prob.range_bins.x0,
                    prob.range_bins.bw, prob.range_bins.nb, prob.lat_y0, prob.lat_z0, prob.lat_bw,
                    prob.lat_nb, prob.vac_bins.x0, prob.vac_bins.bw, prob.vac_bins.nb,
                    int(per_history))
        return pk, o, keep

    def autotune(self, prob: Problem, sample: int = 2000, reps: int = 1, arith: int = MT_FP64):
        """Pick the fastest launch shape for a launch of all of `prob`, timed on
        its first `sample` histories (kept in the context for later runs); when
        `prob` holds at least twice `sample` histories, the finalists are timed
        again on up to four times `sample`.  Returns (Tuning, collisions/s
        projected for the whole launch)."""
        pk, o, keep = self._launch_args(prob, False, arith)
        best = Tuning()
        rate = C.c_double()
        _check(lib().mt_autotune(self.h, C.byref(pk), C.byref(o), int(sample), int(reps),
                                 C.byref(best), C.byref(rate)), "mt_autotune")
        del keep
        return best, rate.value

    def run(self, prob: Problem, per_history: bool = False, arith: int = MT_FP64, log=None,
            noalloc: bool = False) -> dict:
        """Full cascades for every history of `prob` (sets the target if needed).
        `log` (optional): dict(mask=MT_LOG_*, max_records=n, vacmap_Z=(z1, z2, z3));
        the result then holds "records" (RECORD array), "n_records", "vacmap" and
        "transport" (n_mat x 93 net atom moves per layer and Z, MT_LOG_TRANSPORT).
        `noalloc`: through mt_run_cascades_noalloc (the device's static scratch;
        no per-history records or streams)."""
        pk, o, keep = self._launch_args(prob, per_history, arith)
        out = Tallies()
        hists = {"range_hist": np.zeros(prob.range_bins.nb + 2, np.int64),
                 "lat_y_hist": np.zeros(prob.lat_nb + 2, np.int64),
                 "lat_z_hist": np.zeros(prob.lat_nb + 2, np.int64),
                 "vac_hist": np.zeros(prob.vac_bins.nb + 2, np.int64),
                 "repl_hist": np.zeros(prob.vac_bins.nb + 2, np.int64)}
        for k, a in hists.items():
            setattr(out, k, _p(a, lp))
        per = {}
        if per_history:
            n = prob.n_hist
            per = {"h_vac": np.zeros(n, np.int32), "h_coll": np.zeros(n, np.int64),
                   "h_flights": np.zeros(n, np.int32), "h_x": np.zeros(n), "h_y": np.zeros(n),
                   "h_z": np.zeros(n), "h_state": np.zeros(n, np.int32), "h_Eel": np.zeros(n)}
            for k, a in per.items():
                setattr(out, k, _p(a, {np.int32: ip, np.int64: lp, np.float64: dp}[a.dtype.type]))
        if noalloc:
            if log is not None or per_history:
                raise ValueError("the no-allocation run keeps no per-history records or streams")
            _check(lib().mt_run_cascades_noalloc(self.h, C.byref(pk), C.byref(o), C.byref(out)),
                   "mt_run_cascades_noalloc")
        elif log is None:
            _check(lib().mt_run_cascades(self.h, C.byref(pk), C.byref(o), C.byref(out)),
                   "mt_run_cascades")
        else:
            lg = LogSpec()
            lg.mask = int(log.get("mask", 0))
            recs = np.zeros(int(log.get("max_records", 0)), RECORD)
            lg.records = recs.ctypes.data_as(C.c_void_p) if len(recs) else None
            lg.max_records = len(recs)
            vz = log.get("vacmap_Z", (0, 0, 0))
            for i in range(3):
                lg.vacmap_Z[i] = int(vz[i])
            vmap = np.zeros((20, 20, 3), np.int32)
            lg.vacmap = _p(vmap, ip)
            trans = np.zeros((len(prob.materials), 93), np.int64)
            lg.transport = _p(trans, lp)
            _check(lib().mt_run_cascades_log(self.h, C.byref(pk), C.byref(o), C.byref(lg), C.byref(out)),
                   "mt_run_cascades_log")
        res = {k: getattr(out, k) for k in COUNTERS + ENERGIES}
        if log is not None:
            res.update(records=recs[:min(lg.n_records, len(recs))], n_records=lg.n_records, vacmap=vmap,
                       transport=trans)
        res.update(range_n=out.range_n, range_sum=out.range_sum, range_sum2=out.range_sum2,
                   kernel_ms=out.kernel_ms, grid=out.grid, block=out.block, waves=out.waves,
                   queue_peak=out.queue_peak, queue_ops=out.queue_ops, surrogate=out.surrogate,
                   coll_primary=out.coll_primary, sclk_mhz=out.sclk_mhz, half_ms=out.half_ms,
                   flights_at_half=out.flights_at_half, admit_ms=out.admit_ms,
                   flights_at_admit=out.flights_at_admit)
        res.update(hists)
        res.update(per)
        del keep
        return res

    def stopping(self, Z1: int, m1: float, mat: int, E: np.ndarray, arith: int = MT_FP64) -> np.ndarray:
        """getrstop of ion (Z1, m1) in material `mat` at energies E, by the
        device function of precision mode `arith`."""
        E = np.ascontiguousarray(E, np.float64)
        out = np.zeros_like(E)
        if arith == MT_FP64:
            rc = lib().mt_stopping_batch(self.h, Z1, m1, mat, _p(E, dp), len(E), _p(out, dp))
        else:
            rc = lib().mt_stopping_batch_arith(self.h, arith, Z1, m1, mat, _p(E, dp), len(E), _p(out, dp))
        _check(rc, "mt_stopping_batch")
        return out

    def scatter(self, eps: np.ndarray, b: np.ndarray, potential: int = 0, arith: int = MT_FP64):
        """(s2, cos theta, sin theta, iterations) by the device scattering of
        precision mode `arith` (iterations -1: the surrogate answered)."""
        eps = np.ascontiguousarray(eps, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        n = len(eps)
        s2, ct, st = np.zeros(n), np.zeros(n), np.zeros(n)
        it = np.zeros(n, np.int32)
        if arith == MT_FP64:
            rc = lib().mt_scatter_batch(self.h, potential, _p(eps, dp), _p(b, dp), n, _p(s2, dp),
                                        _p(ct, dp), _p(st, dp), _p(it, ip))
        else:
            rc = lib().mt_scatter_batch_arith(self.h, arith, potential, _p(eps, dp), _p(b, dp), n,
                                              _p(s2, dp), _p(ct, dp), _p(st, dp), _p(it, ip))
        _check(rc, "mt_scatter_batch")
        return s2, ct, st, it

    # ---------------------------------------------------------- surrogate
    @staticmethod
    def surrogate_opts(prob: Problem, E_min: float = 3.0, eps_max: float = 1.0, n_train=4_000_000, epochs=60,
                       batch=4096, lr=4e-3, seed=7, bf16=False, inputs=3, sampling=0,
                       eps_min=1e-4) -> SurrogateOpts:
        """Options for the cascades of `prob`: its primary species (target atoms
        are added by the library) and flight energies up to its largest primary
        energy.  inputs=2, sampling=1 with eps_min=1e-4, eps_max=1e4 is the
        2 -> 64 -> 64 -> 1 network on a (log eps, b) rectangle.  The returned
        object keeps its arrays alive."""
        p = prob.primaries
        Zs, ms = [], []
        for Z, m in zip(np.atleast_1d(p.Z), np.atleast_1d(p.m)):
            if (int(Z), float(m)) not in zip(Zs, ms):
                Zs.append(int(Z))
                ms.append(float(m))
        Za = np.array(Zs, np.int32)
        ma = np.array(ms, np.float64)
        o = SurrogateOpts(len(Zs), _p(Za, ip), _p(ma, dp), float(E_min), float(np.max(p.E)), float(eps_max),
                          int(n_train), int(epochs), int(batch), float(lr), int(seed) & 0xFFFFFFFF, int(bf16),
                          int(inputs), int(sampling), float(eps_min))
        o._keep = (Za, ma)
        return o

    @staticmethod
    def surrogate_init(seed: int = 1) -> np.ndarray:
        """Glorot-uniform weights, zero biases (flat layout of mt_mlp.h)."""
        rng = np.random.default_rng(seed)
        p = np.zeros(N_SURROGATE_PARAMS, np.float32)
        def glorot(n_in, n_out):
            a = math.sqrt(6.0 / (n_in + n_out))
            return rng.uniform(-a, a, n_in * n_out).astype(np.float32)
        p[0:192] = glorot(3, 64)
        p[256:256 + 4096] = glorot(64, 64)
        p[4416:4480] = glorot(64, 1)
        return p

    def surrogate_norm(self, opts: SurrogateOpts) -> dict:
        """Input / output scaling and domain of the surrogate for these options."""
        out = np.zeros(len(SURROGATE_NORM_KEYS), np.float32)
        _check(lib().mt_surrogate_norm(self.h, C.byref(opts), _p(out, C.POINTER(C.c_float))), "mt_surrogate_norm")
        return dict(zip(SURROGATE_NORM_KEYS, (float(x) for x in out)))

    def surrogate_train(self, opts: SurrogateOpts, params: np.ndarray):
        params = np.ascontiguousarray(params, np.float32).copy()
        loss = np.zeros(opts.epochs)
        _check(lib().mt_surrogate_train(self.h, C.byref(opts), _p(params, C.POINTER(C.c_float)),
                                        _p(loss, dp)), "mt_surrogate_train")
        return params, loss

    def surrogate_install(self, opts: SurrogateOpts, params: np.ndarray):
        params = np.ascontiguousarray(params, np.float32)
        _check(lib().mt_surrogate_install(self.h, C.byref(opts), _p(params, C.POINTER(C.c_float))),
               "mt_surrogate_install")

    def surrogate_eval(self, eps, b) -> np.ndarray:
        eps = np.ascontiguousarray(eps, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        out = np.zeros(len(eps))
        _check(lib().mt_surrogate_eval(self.h, _p(eps, dp), _p(b, dp), len(eps), _p(out, dp)),
               "mt_surrogate_eval")
        return out

    ABLATE_MODES = {0: "magic_fp32", 1: "mlp_fp32", 2: "mlp_bf16_dot", 3: "mlp_bf16_mfma"}

    def surrogate_ablate(self, mode: int, eps, b, reps: int = 5):
        """s2 at (eps, b) and the mean kernel time of one pass: mode 0 the FP32
        MAGIC solve, 1-3 the installed network with layer 2 in FP32 FMAs, BF16
        dot products or BF16 matrix-core products (no domain check)."""
        eps = np.ascontiguousarray(eps, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        out = np.zeros(len(eps))
        ms = C.c_double(0.0)
        _check(lib().mt_surrogate_ablate(self.h, int(mode), _p(eps, dp), _p(b, dp), len(eps), int(reps),
                                         _p(out, dp), C.byref(ms)), "mt_surrogate_ablate")
        return out, ms.value

    def surrogate_grad(self, opts: SurrogateOpts, params, eps, b, ln_s2):
        params = np.ascontiguousarray(params, np.float32)
        eps = np.ascontiguousarray(eps, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        ln_s2 = np.ascontiguousarray(ln_s2, np.float64)
        g = np.zeros(N_SURROGATE_PARAMS, np.float32)
        loss = C.c_double()
        _check(lib().mt_surrogate_grad(self.h, C.byref(opts), _p(params, C.POINTER(C.c_float)),
                                       _p(eps, dp), _p(b, dp), _p(ln_s2, dp), len(eps),
                                       _p(g, C.POINTER(C.c_float)), C.byref(loss)), "mt_surrogate_grad")
        return g, loss.value

    def trim_flight(self, ion: Ion, u: np.ndarray, max_events: int = 1 << 16):
        """One TrimBase::trim() flight driven by the uniforms `u`.  Returns
        (draws used, events) or raises MtError when `u` ran out."""
        u = np.ascontiguousarray(u, np.float64)
        used = C.c_longlong()
        nev = C.c_int()
        ev = (Event * max_events)()
        _check(lib().mt_trim_flight(self.h, C.byref(ion), _p(u, dp), len(u), C.byref(used), ev,
                                    max_events, C.byref(nev)), "mt_trim_flight")
        return used.value, ev[:nev.value]
