"""ctypes binding of libsouffle_gpu.so (C ABI in include/souffle_gpu.h).

Thin convenience layer for Python callers: it converts numpy arrays to the library's plain
pointers and raises on error codes. There is no fallback: if the HIP library is not built or
no GPU is present, loading or the first device call fails loudly.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsouffle_gpu.so")
REPO = os.path.dirname(_HERE)


class SfCli(ctypes.Structure):
    _fields_ = [("C_uM", ctypes.c_double), ("c_mM", ctypes.c_double), ("ks", ctypes.c_double),
                ("kb", ctypes.c_double), ("elj_att", ctypes.c_double), ("ecut_c", ctypes.c_double),
                ("dt", ctypes.c_double), ("K", ctypes.c_double), ("Q", ctypes.c_double),
                ("damp", ctypes.c_double), ("n_subunits", ctypes.c_int), ("q_chain", ctypes.c_int),
                ("brownian", ctypes.c_int)]


class SfParams(ctypes.Structure):
    _fields_ = [("box", ctypes.c_double), ("ecut", ctypes.c_double), ("lb", ctypes.c_double),
                ("kappa", ctypes.c_double), ("ecut_el", ctypes.c_double), ("ks", ctypes.c_double),
                ("kb", ctypes.c_double), ("elj_att", ctypes.c_double), ("T", ctypes.c_double),
                ("dt", ctypes.c_double), ("damp", ctypes.c_double), ("Q", ctypes.c_double),
                ("q_chain", ctypes.c_int), ("brownian", ctypes.c_int), ("skin", ctypes.c_double),
                ("pair_fp32", ctypes.c_int), ("bd_seed", ctypes.c_uint32), ("far_rf", ctypes.c_double)]


class SfStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("rebuilds", ctypes.c_int64), ("lj_entries", ctypes.c_int64),
                ("q_entries", ctypes.c_int64), ("lj_capacity", ctypes.c_int32), ("q_capacity", ctypes.c_int32),
                ("n_beads", ctypes.c_int32), ("n_subunits", ctypes.c_int32), ("n_charged", ctypes.c_int32),
                ("n_replicas", ctypes.c_int32), ("k_integrate_ms", ctypes.c_double),
                ("k_rebuild_ms", ctypes.c_double),
                ("k_force_ms", ctypes.c_double), ("k_reduce_ms", ctypes.c_double),
                ("k_timed_steps", ctypes.c_double), ("vcom", ctypes.c_double), ("k_far_ms", ctypes.c_double),
                ("graph_launches", ctypes.c_int64), ("graph_steps", ctypes.c_int32)]


class SfResume(ctypes.Structure):
    _fields_ = [("step", ctypes.c_int64), ("q_chain", ctypes.c_int32), ("rng", ctypes.c_uint32),
                ("ke", ctypes.c_double), ("xi", ctypes.c_double * 8), ("eta", ctypes.c_double * 8)]


class SfTuneOpts(ctypes.Structure):
    _fields_ = [("trial_steps", ctypes.c_int), ("n_skins", ctypes.c_int), ("skins", ctypes.POINTER(ctypes.c_double)),
                ("n_tiles", ctypes.c_int), ("tiles", ctypes.POINTER(ctypes.c_int))]


class SfTuneResult(ctypes.Structure):
    _fields_ = [("skin", ctypes.c_double), ("tile", ctypes.c_int), ("ms_per_step", ctypes.c_double),
                ("n_trials", ctypes.c_int)]


class SfTrainOpts(ctypes.Structure):
    _fields_ = [("n_samples", ctypes.c_int), ("epochs", ctypes.c_int), ("batch", ctypes.c_int),
                ("lr", ctypes.c_double), ("val_frac", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("c_lo_mM", ctypes.c_double), ("c_hi_mM", ctypes.c_double)]


class SfTrainReport(ctypes.Structure):
    _fields_ = [("train_loss", ctypes.c_double), ("val_loss", ctypes.c_double),
                ("val_rel_err_force", ctypes.c_double), ("val_rel_err_torque", ctypes.c_double),
                ("train_rel_err_force", ctypes.c_double), ("seconds", ctypes.c_double), ("steps", ctypes.c_int)]


class SfFarStats(ctypes.Structure):
    _fields_ = [("audits", ctypes.c_int64), ("last_err", ctypes.c_double), ("max_err", ctypes.c_double),
                ("using_model", ctypes.c_int), ("fell_back", ctypes.c_int), ("last_err_force", ctypes.c_double),
                ("max_err_force", ctypes.c_double)]


# every entry point of include/souffle_gpu.h: name -> (restype, argtypes)
_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_FP = ctypes.POINTER(ctypes.c_float)
SIGNATURES = {
    "sf_template_load": (_P, [ctypes.c_char_p]),
    "sf_template_free": (None, [_P]),
    "sf_template_info": (ctypes.c_int, [_P, _I, _I, _I, _I, _I, _I, _D, _D]),
    "sf_template_hinges": (ctypes.c_int, [_P, _I, ctypes.c_int]),
    "sf_template_extent": (ctypes.c_int, [_P, _D, _D, _D]),
    "sf_params_derive": (ctypes.c_int, [_P, ctypes.POINTER(SfCli), ctypes.POINTER(SfParams)]),
    "sf_make_state": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _D, _D]),
    "sf_create": (_P, [_P, ctypes.c_int, _I, ctypes.POINTER(SfParams)]),
    "sf_destroy": (None, [_P]),
    "sf_set_state": (ctypes.c_int, [_P, _D, _D]),
    "sf_get_state": (ctypes.c_int, [_P, _D, _D]),
    "sf_get_thermostat": (ctypes.c_int, [_P, ctypes.c_int, _D, _D]),
    "sf_get_resume": (ctypes.c_int, [_P, ctypes.POINTER(SfResume), _D]),
    "sf_set_state_resume": (ctypes.c_int, [_P, _D, _D, ctypes.POINTER(SfResume), _D]),
    "sf_set_positions": (ctypes.c_int, [_P, _D]),
    "sf_forces_ex": (ctypes.c_int, [_P, ctypes.c_int, _D, _D, _D, _D, _D]),
    "sf_forces": (ctypes.c_int, [_P, _D, _D, _D, _D, _D]),
    "sf_energies": (ctypes.c_int, [_P, _D, _D]),
    "sf_step": (ctypes.c_int, [_P, ctypes.c_int]),
    "sf_set_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "sf_set_graph_steps": (ctypes.c_int, [_P, ctypes.c_int]),
    "sf_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(SfStats)]),
    "sf_bd_uniforms": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int64, _D]),
    "sf_set_skin": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double]),
    "sf_get_skin": (ctypes.c_int, [_P, ctypes.c_int, _D]),
    "sf_set_tile": (ctypes.c_int, [_P, ctypes.c_int]),
    "sf_autotune": (ctypes.c_int, [_P, ctypes.POINTER(SfTuneOpts), ctypes.POINTER(SfTuneResult), _D]),
    "sf_template_frame": (ctypes.c_int, [_P, _I, _I, _D]),
    "sf_far_model_create": (_P, [_P, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_uint64]),
    "sf_far_model_destroy": (None, [_P]),
    "sf_far_model_info": (ctypes.c_int, [_P, _D, _D, _D, _I]),
    "sf_far_model_params": (ctypes.c_int, [_P, _FP, _FP, _FP, _FP]),
    "sf_far_model_samples": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                            _FP, _FP, _D]),
    "sf_far_model_grad": (ctypes.c_int, [_P, ctypes.c_int, _FP, _FP, _FP, _D]),
    "sf_far_model_infer": (ctypes.c_int, [_P, ctypes.c_int, _FP, _FP, ctypes.c_int]),
    "sf_far_model_train": (ctypes.c_int, [_P, ctypes.POINTER(SfTrainOpts), ctypes.POINTER(SfTrainReport)]),
    "sf_set_far_model": (ctypes.c_int, [_P, _P, ctypes.c_double]),
    "sf_far_audit": (ctypes.c_int, [_P, _D]),
    "sf_far_drop": (ctypes.c_int, [_P, ctypes.c_int]),
    "sf_get_far_stats": (ctypes.c_int, [_P, ctypes.POINTER(SfFarStats)]),
    "sf_oligomers": (ctypes.c_int, [_P, ctypes.c_double, _I, _I]),
    "sf_oligomers_noalloc": (ctypes.c_int, [_P, ctypes.c_double, _I, _I]),
    "sf_big_beads_place": (ctypes.c_int, [_P, ctypes.c_int, _D, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_uint, _D]),
    "sf_set_big_beads": (ctypes.c_int, [_P, ctypes.c_int, _D, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "sf_big_energies": (ctypes.c_int, [_P, _D]),
    "sf_last_error": (ctypes.c_char_p, []),
}

_lib = None


def load(path=LIB_PATH):
    """Load the HIP library (raises OSError if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise OSError(f"{path} is not built: run __graft_entry__.build() or make -C capsid-souffle_amd")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


class SfError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


SF_E_NONFINITE = -5


def _err(rc, what):
    if rc < 0:
        raise SfError(f"{what} failed ({rc}): {_lib.sf_last_error().decode()}", rc)
    return rc


def _dp(a):
    if a is None:
        return None
    return a.ctypes.data_as(_D)


class Template:
    def __init__(self, path):
        lib = load()
        self.path = path
        self.ptr = lib.sf_template_load(path.encode())
        if not self.ptr:
            raise SfError(lib.sf_last_error().decode())
        ints = [ctypes.c_int() for _ in range(6)]
        dbl = [ctypes.c_double() for _ in range(2)]
        _err(lib.sf_template_info(self.ptr, *[ctypes.byref(x) for x in ints], *[ctypes.byref(x) for x in dbl]),
             "sf_template_info")
        (self.n_beads, self.n_edges, self.n_faces, self.n_hinges, self.n_charged,
         self.centre) = [x.value for x in ints]
        self.si_sigma, self.si_mass = [x.value for x in dbl]
        ext = [ctypes.c_double() for _ in range(3)]
        _err(lib.sf_template_extent(self.ptr, *[ctypes.byref(x) for x in ext]), "sf_template_extent")
        self.diameter, self.r_all, self.r_charged = [x.value for x in ext]

    def frame(self):
        """(fp1, fp2, qbody [n_charged][3]) of the subunit body frame."""
        a, b = ctypes.c_int(), ctypes.c_int()
        qb = np.zeros((self.n_charged, 3))
        _err(load().sf_template_frame(self.ptr, ctypes.byref(a), ctypes.byref(b), _dp(qb)), "sf_template_frame")
        return a.value, b.value, qb

    def hinges(self):
        out = np.zeros((self.n_hinges, 4), dtype=np.int32)
        load().sf_template_hinges(self.ptr, out.ctypes.data_as(_I), self.n_hinges)
        return out

    def params(self, **cli):
        """sf_params from the reference's flags (C_uM, c_mM, ks, kb, elj_att, ecut_c, dt, K, Q,
        damp, n_subunits, q_chain, brownian); defaults are the reference's defaults."""
        c = SfCli(C_uM=200.0, c_mM=500.0, ks=50.0, kb=20.0, elj_att=2.0, ecut_c=12.0, dt=0.004,
                  K=298.0, Q=10.0, damp=100.0, n_subunits=64, q_chain=5, brownian=0)
        for k, v in cli.items():
            setattr(c, k, v)
        p = SfParams()
        _err(load().sf_params_derive(self.ptr, ctypes.byref(c), ctypes.byref(p)), "sf_params_derive")
        return p

    def make_state(self, n_subunits, box, T, seed):
        n = n_subunits * self.n_beads
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        _err(load().sf_make_state(self.ptr, n_subunits, box, T, seed, _dp(pos), _dp(vel)), "sf_make_state")
        return pos, vel

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.sf_template_free(self.ptr)
            self.ptr = None


class System:
    """A device-resident system (or batch of replicas) on the current GPU."""

    def __init__(self, template, n_subunits, params):
        lib = load()
        if isinstance(n_subunits, int):
            n_subunits = [n_subunits]
        if isinstance(params, SfParams):
            params = [params] * len(n_subunits)
        self.template = template
        self.n_subunits = list(n_subunits)
        self.R = len(n_subunits)
        ns = (ctypes.c_int * self.R)(*n_subunits)
        pa = (SfParams * self.R)(*params)
        self.ptr = lib.sf_create(template.ptr, self.R, ns, pa)
        if not self.ptr:
            raise SfError(lib.sf_last_error().decode())
        self.n_beads = sum(n_subunits) * template.n_beads

    def set_state(self, pos, vel=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64)
        assert pos.shape == (self.n_beads, 3)
        _err(_lib.sf_set_state(self.ptr, _dp(pos), _dp(vel)), "sf_set_state")

    def get_state(self):
        pos = np.zeros((self.n_beads, 3))
        vel = np.zeros((self.n_beads, 3))
        _err(_lib.sf_get_state(self.ptr, _dp(pos), _dp(vel)), "sf_get_state")
        return pos, vel

    def get_resume(self, with_force=True):
        """(resume records [R] as dicts, integrator force [N,3] or None): what sf_set_state_resume
        needs besides positions and velocities to continue bit for bit."""
        recs = (SfResume * self.R)()
        f = np.zeros((self.n_beads, 3)) if with_force else None
        _err(_lib.sf_get_resume(self.ptr, recs, _dp(f)), "sf_get_resume")
        out = []
        for r in recs:
            q = r.q_chain
            out.append(dict(step=r.step, q_chain=q, rng=r.rng, ke=r.ke, xi=list(r.xi)[:q], eta=list(r.eta)[:q]))
        return out, f

    def set_state_resume(self, pos, vel, records, tforce=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        assert pos.shape == (self.n_beads, 3) and vel.shape == (self.n_beads, 3) and len(records) == self.R
        recs = (SfResume * self.R)()
        for r, d in zip(recs, records):
            r.step, r.q_chain, r.rng, r.ke = int(d["step"]), int(d["q_chain"]), int(d["rng"]), float(d["ke"])
            for j, (x, e) in enumerate(zip(d["xi"], d["eta"])):
                r.xi[j], r.eta[j] = x, e
        tf = None if tforce is None else np.ascontiguousarray(tforce, dtype=np.float64)
        _err(_lib.sf_set_state_resume(self.ptr, _dp(pos), _dp(vel), recs, _dp(tf)), "sf_set_state_resume")

    def set_positions(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        assert pos.shape == (self.n_beads, 3)
        _err(_lib.sf_set_positions(self.ptr, _dp(pos)), "sf_set_positions")

    def forces_ex(self, force_rebuild):
        """sf_forces_ex: (total, stretch, bend, yukawa, lj) with the reference's updatePairlist flag."""
        out = [np.zeros((self.n_beads, 3)) for _ in range(5)]
        _err(_lib.sf_forces_ex(self.ptr, 1 if force_rebuild else 0, *[_dp(a) for a in out]), "sf_forces_ex")
        return out

    def thermostat(self, replica=0):
        xi = np.zeros(16)
        eta = np.zeros(16)
        q = _err(_lib.sf_get_thermostat(self.ptr, replica, _dp(xi), _dp(eta)), "sf_get_thermostat")
        return xi[:q], eta[:q]

    def forces(self):
        """(total, stretch, bend, yukawa, lj), each [N,3]."""
        out = [np.zeros((self.n_beads, 3)) for _ in range(5)]
        _err(_lib.sf_forces(self.ptr, *[_dp(a) for a in out]), "sf_forces")
        return out

    def energies(self, per_bead=False):
        tot = np.zeros((self.R, 7))
        pb = np.zeros((self.n_beads, 5)) if per_bead else None
        _err(_lib.sf_energies(self.ptr, _dp(tot), _dp(pb)), "sf_energies")
        return (tot, pb) if per_bead else tot

    def step(self, n):
        _err(_lib.sf_step(self.ptr, int(n)), "sf_step")

    def set_timing(self, on):
        _err(_lib.sf_set_timing(self.ptr, 1 if on else 0), "sf_set_timing")

    def set_graph_steps(self, steps):
        """Steps per HIP graph launch of step() (0: the default 100; at most 20000)."""
        _err(_lib.sf_set_graph_steps(self.ptr, int(steps)), "sf_set_graph_steps")

    def stats(self):
        s = SfStats()
        _err(_lib.sf_get_stats(self.ptr, ctypes.byref(s)), "sf_get_stats")
        return {k: getattr(s, k) for k, _ in SfStats._fields_}

    def bd_uniforms(self, replica, step):
        nb = self.n_subunits[replica] * self.template.n_beads
        out = np.zeros((nb, 3))
        _err(_lib.sf_bd_uniforms(self.ptr, replica, int(step), _dp(out)), "sf_bd_uniforms")
        return out

    def set_skin(self, skin, replica=-1):
        _err(_lib.sf_set_skin(self.ptr, int(replica), float(skin)), "sf_set_skin")

    def skin(self, replica=0):
        v = ctypes.c_double()
        _err(_lib.sf_get_skin(self.ptr, int(replica), ctypes.byref(v)), "sf_get_skin")
        return v.value

    def set_tile(self, tile):
        _err(_lib.sf_set_tile(self.ptr, int(tile)), "sf_set_tile")

    def autotune(self, skins=None, tiles=None, trial_steps=0):
        """Runs the library's autotuner; returns (result dict, trial ms/step [n_skins][n_tiles])."""
        o = SfTuneOpts(trial_steps=int(trial_steps))
        sk = np.ascontiguousarray(skins if skins is not None else [0.2, 0.3, 0.5, 0.8, 1.2], dtype=np.float64)
        tl = np.ascontiguousarray(tiles if tiles is not None else [128, 256, 384, 512], dtype=np.int32)
        o.n_skins, o.skins = len(sk), sk.ctypes.data_as(_D)
        o.n_tiles, o.tiles = len(tl), tl.ctypes.data_as(_I)
        res = SfTuneResult()
        ms = np.zeros((len(sk), len(tl)))
        _err(_lib.sf_autotune(self.ptr, ctypes.byref(o), ctypes.byref(res), _dp(ms)), "sf_autotune")
        return {k: getattr(res, k) for k, _ in SfTuneResult._fields_}, ms

    def set_far_model(self, model, tol=1e-3):
        _err(_lib.sf_set_far_model(self.ptr, model.ptr if model is not None else None, float(tol)), "sf_set_far_model")

    def far_drop(self, drop=True):
        _err(_lib.sf_far_drop(self.ptr, 1 if drop else 0), "sf_far_drop")

    def far_audit(self):
        v = ctypes.c_double()
        _err(_lib.sf_far_audit(self.ptr, ctypes.byref(v)), "sf_far_audit")
        return v.value

    def far_stats(self):
        s = SfFarStats()
        _err(_lib.sf_get_far_stats(self.ptr, ctypes.byref(s)), "sf_get_far_stats")
        return {k: getattr(s, k) for k, _ in SfFarStats._fields_}

    def oligomers(self, contact=2.0):
        """Oligomers at the current positions, per replica: a list of (label, mass_spectrum)
        int32 arrays of that replica's subunit count (label = the smallest subunit of each
        subunit's oligomer; mass_spectrum[k - 1] = oligomers of k subunits)."""
        ns = sum(self.n_subunits)
        lab = np.zeros(ns, dtype=np.int32)
        ms = np.zeros(ns, dtype=np.int32)
        _err(_lib.sf_oligomers(self.ptr, float(contact), lab.ctypes.data_as(_I), ms.ctypes.data_as(_I)),
             "sf_oligomers")
        out, s0 = [], 0
        for n in self.n_subunits:
            out.append((lab[s0:s0 + n].copy(), ms[s0:s0 + n].copy()))
            s0 += n
        return out

    def oligomers_noalloc(self, contact=2.0):
        """As oligomers(), through the allocation-free entry point (fixed device buffers, at most
        262,144 subunits per handle); also returns the oligomer count it reports."""
        ns = sum(self.n_subunits)
        lab = np.zeros(ns, dtype=np.int32)
        ms = np.zeros(ns, dtype=np.int32)
        n = _err(_lib.sf_oligomers_noalloc(self.ptr, float(contact), lab.ctypes.data_as(_I), ms.ctypes.data_as(_I)),
                 "sf_oligomers_noalloc")
        out, s0 = [], 0
        for k in self.n_subunits:
            out.append((lab[s0:s0 + k].copy(), ms[s0:s0 + k].copy()))
            s0 += k
        return out, n

    def set_big_beads(self, xyz, sigma=6.0, charge=-78.0, shc=3.5):
        """Immobile encapsulation big beads at xyz[n][3] (defaults: the reference's -a 3, -G -78,
        -H 3.5); applies from the next set_state(). An empty array removes them."""
        xyz = np.ascontiguousarray(np.reshape(xyz, (-1, 3)), dtype=np.float64)
        _err(_lib.sf_set_big_beads(self.ptr, xyz.shape[0], _dp(xyz), float(sigma), float(charge), float(shc)),
             "sf_set_big_beads")

    def big_energies(self):
        """(sum of ce, sum of ne) of the big beads at the current positions."""
        out = np.zeros(2)
        _err(_lib.sf_big_energies(self.ptr, _dp(out)), "sf_big_energies")
        return out

    def close(self):
        if getattr(self, "ptr", None):
            _lib.sf_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()


def place_big_beads(template, pos, box, n, sigma=6.0, max_attempts=1000000):
    """The reference's big-bead placement (srand(0), rand() positions, overlap rejection) among
    the subunit beads pos[N][3]; returns [n][3]."""
    load()
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    ns = pos.shape[0] // template.n_beads
    out = np.zeros((n, 3))
    _err(_lib.sf_big_beads_place(template.ptr, ns, _dp(pos), float(box), int(n), float(sigma), int(max_attempts),
                                 _dp(out)), "sf_big_beads_place")
    return out


def _fp(a):
    return None if a is None else a.ctypes.data_as(_FP)


class FarModel:
    """Far-field network of the uncut Yukawa (13 -> 128 -> 128 -> 128 -> 6, SiLU) on the GPU."""

    N_IN, N_OUT = 13, 6

    def __init__(self, template, far_rf=0.0, c_lo=5.0, c_hi=200.0, seed=1):
        lib = load()
        self.ptr = lib.sf_far_model_create(template.ptr, float(far_rf), float(c_lo), float(c_hi), int(seed))
        if not self.ptr:
            raise SfError(lib.sf_last_error().decode())
        rf, k0, k1, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        _err(lib.sf_far_model_info(self.ptr, ctypes.byref(rf), ctypes.byref(k0), ctypes.byref(k1), ctypes.byref(n)),
             "sf_far_model_info")
        self.far_rf, self.kappa_min, self.kappa_max, self.n_params = rf.value, k0.value, k1.value, n.value

    def params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        norm = np.zeros(12, dtype=np.float32)
        _err(_lib.sf_far_model_params(self.ptr, _fp(p), None, _fp(norm), None), "sf_far_model_params")
        return p, norm[:6].copy(), norm[6:].copy()

    def set_params(self, p=None, mu=None, sd=None):
        p = None if p is None else np.ascontiguousarray(p, dtype=np.float32)
        norm = None
        if mu is not None:
            norm = np.ascontiguousarray(np.concatenate([mu, sd]), dtype=np.float32)
        _err(_lib.sf_far_model_params(self.ptr, None, _fp(p), None, _fp(norm)), "sf_far_model_params")

    def samples(self, n, seed, c_lo, c_hi):
        x = np.zeros((n, 13), dtype=np.float32)
        y = np.zeros((n, 6), dtype=np.float32)
        pose = np.zeros((n, 16))
        _err(_lib.sf_far_model_samples(self.ptr, int(n), int(seed), float(c_lo), float(c_hi), _fp(x), _fp(y), _dp(pose)),
             "sf_far_model_samples")
        return x, y, pose

    def grad(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = ctypes.c_double()
        _err(_lib.sf_far_model_grad(self.ptr, len(x), _fp(x), _fp(y), _fp(g), ctypes.byref(loss)), "sf_far_model_grad")
        return loss.value, g

    def infer(self, x, bf16_mfma=True):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((len(x), 6), dtype=np.float32)
        _err(_lib.sf_far_model_infer(self.ptr, len(x), _fp(x), _fp(y), 1 if bf16_mfma else 0), "sf_far_model_infer")
        return y

    def train(self, n_samples=1000000, epochs=20, batch=4096, lr=2e-3, val_frac=0.05, seed=12345, c_lo=0.0, c_hi=0.0):
        o = SfTrainOpts(n_samples=n_samples, epochs=epochs, batch=batch, lr=lr, val_frac=val_frac, seed=seed,
                        c_lo_mM=c_lo, c_hi_mM=c_hi)
        r = SfTrainReport()
        _err(_lib.sf_far_model_train(self.ptr, ctypes.byref(o), ctypes.byref(r)), "sf_far_model_train")
        return {k: getattr(r, k) for k, _ in SfTrainReport._fields_}

    def close(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.sf_far_model_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()
