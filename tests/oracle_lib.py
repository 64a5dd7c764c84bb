"""ctypes wrapper of the CPU oracle (oracle/liboracle.so). Test infrastructure only."""
import ctypes
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
TEMPLATE = os.path.join(REPO, "tests", "golden", "infiles", "trimer_MVM")       # data/model copy, charged
TEMPLATE_BIN = os.path.join(REPO, "tests", "golden", "infiles", "trimer_MVM_bin")  # bin/infiles copy, q = 0

_D = ctypes.POINTER(ctypes.c_double)
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.check_call(["make", "-C", os.path.join(REPO, "oracle")])
        lib = ctypes.CDLL(ORACLE_SO)
        lib.orc_create.restype = ctypes.c_void_p
        lib.orc_create.argtypes = [ctypes.c_char_p, _D, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        lib.orc_destroy.argtypes = [ctypes.c_void_p]
        lib.orc_params.argtypes = [ctypes.c_void_p, _D]
        lib.orc_set_state.argtypes = [ctypes.c_void_p, _D, _D]
        lib.orc_get_state.argtypes = [ctypes.c_void_p, _D, _D, _D, _D]
        lib.orc_set_thermostat.argtypes = [ctypes.c_void_p, _D, _D, ctypes.c_int, ctypes.c_double]
        lib.orc_set_list_capacity.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.orc_forces.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _D, _D, _D, _D]
        lib.orc_energies.argtypes = [ctypes.c_void_p, _D, _D]
        lib.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int]
        lib.orc_time_steps.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_int]
        lib.orc_time_steps.restype = ctypes.c_double
        lib.orc_list_entries.argtypes = [ctypes.c_void_p]
        lib.orc_list_entries.restype = ctypes.c_longlong
        lib.orc_neighbours.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        lib.orc_hinges.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        lib.orc_bd_uniforms.argtypes = [ctypes.c_longlong, ctypes.c_int, _D]
        lib.orc_oligomers.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_int)]
        lib.orc_set_big_beads.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double]
        lib.orc_big_energies.argtypes = [ctypes.c_void_p, _D]
        lib.orc_bead_props.argtypes = [ctypes.c_void_p, _D, _D]
        lib.orc_place_big_beads.argtypes = [_D, _D, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_uint, _D]
        lib.orc_last_error.restype = ctypes.c_char_p
        _lib = lib
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


class Oracle:
    """CPU restatement of the reference path for one system (see oracle/souffle_oracle.cpp)."""

    def __init__(self, n_subunits, template=TEMPLATE, C_uM=800.0, c_mM=100.0, ks=50.0, kb=20.0,
                 elj_att=2.1, ecut_c=12.0, L=7.5, B=20, dt=0.005, K=298.0, Q=10.0, damp=100.0,
                 q_chain=5, brownian=False, long_double=False):
        lib = load()
        cd = (ctypes.c_double * 11)(C_uM, c_mM, ks, kb, elj_att, ecut_c, L, dt, K, Q, damp)
        ci = (ctypes.c_int * 4)(n_subunits, B, q_chain, 1 if brownian else 0)
        self.ptr = lib.orc_create(template.encode(), cd, ci, 1 if long_double else 0)
        if not self.ptr:
            raise RuntimeError(lib.orc_last_error().decode())
        p = np.zeros(13)
        lib.orc_params(self.ptr, _dp(p))
        (self.box, self.ecut, self.lb, self.kappa, self.ecut_el, self.rc_list, self.T, self.si_time) = p[:8]
        self.nlv, self.N, self.Ns, self.np_ = int(p[8]), int(p[9]), int(p[10]), int(p[11])
        self.step_index = 0

    def set_list_capacity(self, nlv):
        """Neighbour-list slots per bead (the reference's fixed capacity is sized for its usual
        densities); call before set_state."""
        if load().orc_set_list_capacity(self.ptr, int(nlv)) != 0:
            raise RuntimeError("bad list capacity")
        self.nlv = min(int(nlv), self.N)

    def centre_local(self):
        p = np.zeros(13)
        load().orc_params(self.ptr, _dp(p))
        return int(p[12])

    def set_state(self, pos, vel=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.zeros_like(pos) if vel is None else np.ascontiguousarray(vel, dtype=np.float64)
        rc = load().orc_set_state(self.ptr, _dp(pos), _dp(vel))
        self.step_index = 0
        if rc:
            raise RuntimeError("oracle neighbour list overflow")

    def set_thermostat(self, xi, eta, particle_ke, step_index=0):
        """Continue another program's NH run: chain xi/eta, the kinetic energy the next step's
        first chain update reads, and the loop index of the next step (list-rebuild schedule)."""
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        if load().orc_set_thermostat(self.ptr, _dp(xi), _dp(eta), len(xi), float(particle_ke)):
            raise RuntimeError("chain length differs from the oracle's")
        self.step_index = step_index

    def get_state(self):
        pos = np.zeros((self.N, 3))
        vel = np.zeros((self.N, 3))
        xi = np.zeros(16)
        eta = np.zeros(16)
        q = load().orc_get_state(self.ptr, _dp(pos), _dp(vel), _dp(xi), _dp(eta))
        return pos, vel, xi[:q], eta[:q]

    def forces(self, rebuild=True):
        out = [np.zeros((self.N, 3)) for _ in range(5)]
        rc = load().orc_forces(self.ptr, 1 if rebuild else 0, *[_dp(a) for a in out])
        if rc:
            raise RuntimeError("oracle neighbour list overflow")
        return out

    def energies(self, per_bead=False):
        tot = np.zeros(7)
        pb = np.zeros((self.N, 5)) if per_bead else None
        load().orc_energies(self.ptr, _dp(tot), _dp(pb))
        return (tot, pb) if per_bead else tot

    def run(self, n):
        rc = load().orc_run(self.ptr, self.step_index, n)
        self.step_index += n
        if rc:
            raise RuntimeError("oracle neighbour list overflow")

    def time_steps(self, n, threads=0):
        t = load().orc_time_steps(self.ptr, self.step_index, n, threads)
        self.step_index += n
        return t

    def neighbours(self, i):
        buf = (ctypes.c_int * self.N)()
        n = load().orc_neighbours(self.ptr, i, buf, self.N)
        return np.array(buf[:n], dtype=np.int64)

    def hinges(self):
        buf = (ctypes.c_int * 4096)()
        n = load().orc_hinges(self.ptr, buf, 1024)
        return np.array(buf[:4 * n], dtype=np.int64).reshape(n, 4)

    def oligomers(self, contact=2.0):
        """(label[Ns], mass_spectrum[Ns], count) of the current positions."""
        lab = np.zeros(self.Ns, dtype=np.int32)
        ms = np.zeros(self.Ns, dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        n = load().orc_oligomers(self.ptr, contact, lab.ctypes.data_as(ip), ms.ctypes.data_as(ip))
        return lab, ms, n

    def bead_props(self):
        """(sigma[N], q[N]) of the subunit beads."""
        sig = np.zeros(self.N)
        q = np.zeros(self.N)
        load().orc_bead_props(self.ptr, _dp(sig), _dp(q))
        return sig, q

    def set_big_beads(self, xyz, sigma, q, shc):
        """Encapsulation big beads (immobile) at xyz[n][3]; effective from the next force call."""
        xyz = np.ascontiguousarray(np.reshape(xyz, (-1, 3)), dtype=np.float64)
        load().orc_set_big_beads(self.ptr, xyz.shape[0], _dp(xyz), sigma, q, shc)

    def big_energies(self):
        """(sum of ce, sum of ne) over the big beads from the last energies() call."""
        out = np.zeros(2)
        load().orc_big_energies(self.ptr, _dp(out))
        return out

    def close(self):
        if getattr(self, "ptr", None):
            load().orc_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()


def bd_uniforms(skip, n):
    """The reference's BD noise stream: std::default_random_engine(69521) through
    std::uniform_real_distribution(-0.5, 0.5), draws skip .. skip + n - 1."""
    out = np.zeros(n)
    load().orc_bd_uniforms(skip, n, _dp(out))
    return out


def place_big_beads(pos, sig, box, n, sigma, max_attempts=1000000):
    """The reference's big-bead placement (srand(0), rand() positions, overlap rejection);
    None when a bead needs max_attempts tries."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    sig = np.ascontiguousarray(sig, dtype=np.float64)
    out = np.zeros((n, 3))
    rc = load().orc_place_big_beads(_dp(pos), _dp(sig), sig.shape[0], box, n, sigma, max_attempts, _dp(out))
    return None if rc else out


def rel_norm(a, b):
    """Norm-wise relative difference ||a - b|| / ||b||."""
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
