"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker, never the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)
up = C.POINTER(C.c_uint32)
bp = C.POINTER(C.c_uint8)


def left_pack(beads, nxt, prv, worm):
    """Path::leftPack (src/path.cpp:145-189) on copies of the arrays."""
    b = np.array(beads, dtype=np.float64, order="C")
    n = np.array(nxt, dtype=np.int32, order="C")
    p = np.array(prv, dtype=np.int32, order="C")
    w = np.array(worm, dtype=np.uint32, order="C")
    M, N = w.shape
    lib().orc_left_pack(M, N, _p(b, dp), _p(n, ip), _p(p, ip), w.ctypes.data_as(C.POINTER(C.c_uint32)))
    return b, n, p, w


ENERGY_COLUMNS = ("K", "V", "V_ext", "V_int", "E", "E_mu", "K/N", "V/N", "E/N")
VIRIAL_COLUMNS = ("K_op", "K_cv", "V_op", "V_cv", "E", "E_mu", "K_op/N", "K_cv/N", "V_op/N", "V_cv/N", "E/N",
                  "EEcv*Beta^2", "Ecv*Beta", "dEdB", "CvCov1", "CvCov2", "CvCov3", "E_th", "P")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "orc_mt_create": (C.c_void_p, [C.c_uint32]),
            "orc_mt_destroy": (None, [C.c_void_p]),
            "orc_mt_randint": (C.c_uint32, [C.c_void_p]),
            "orc_mt_randint_n": (C.c_uint32, [C.c_void_p, C.c_uint32]),
            "orc_mt_rand": (C.c_double, [C.c_void_p]),
            "orc_mt_randexc": (C.c_double, [C.c_void_p]),
            "orc_mt_randnorm": (C.c_double, [C.c_void_p, C.c_double, C.c_double]),
            "orc_mt_save": (None, [C.c_void_p, up]),
            "orc_mt_load": (None, [C.c_void_p, up]),
            "orc_aziz_create": (C.c_void_p, [C.c_int, C.c_double, C.c_double, C.c_int]),
            "orc_aziz_destroy": (None, [C.c_void_p]),
            "orc_aziz_valueV": (C.c_double, [C.c_void_p, C.c_double]),
            "orc_aziz_valuedVdr": (C.c_double, [C.c_void_p, C.c_double]),
            "orc_aziz_valued2Vdr2": (C.c_double, [C.c_void_p, C.c_double]),
            "orc_aziz_V": (C.c_double, [C.c_void_p, dp]),
            "orc_aziz_gradV": (None, [C.c_void_p, dp, dp]),
            "orc_aziz_grad2V": (C.c_double, [C.c_void_p, dp]),
            "orc_aziz_tailV": (C.c_double, [C.c_void_p]),
            "orc_aziz_dr": (C.c_double, [C.c_void_p]),
            "orc_aziz_table_length": (C.c_int, [C.c_void_p]),
            "orc_aziz_table": (dp, [C.c_void_p, C.c_int]),
            "orc_aziz_branch_bins": (None, [C.c_void_p, ip]),
            "orc_aziz_params": (None, [C.c_void_p, dp]),
            "orc_state_create": (C.c_void_p, [C.c_double, C.c_int, C.c_int, C.c_double, C.c_double,
                                              C.c_double, C.c_int]),
            "orc_state_destroy": (None, [C.c_void_p]),
            "orc_state_set_beads": (None, [C.c_void_p, dp]),
            "orc_state_set_topology": (None, [C.c_void_p, bp, ip, ip, ip, ip, ip, ip, ip]),
            "orc_state_get_beads": (None, [C.c_void_p, dp]),
            "orc_state_set_action": (None, [C.c_void_p, C.c_int, C.c_double, C.c_double]),
            "orc_state_aziz": (C.c_void_p, [C.c_void_p]),
            "orc_total_action": (C.c_double, [C.c_void_p]),
            "orc_total_action_threads": (C.c_double, [C.c_void_p, C.c_int]),
            "orc_V_slice": (None, [C.c_void_p, C.c_int, dp, ip]),
            "orc_Vnn_bead": (C.c_double, [C.c_void_p, C.c_int, C.c_int]),
            "orc_Vnn_slice": (C.c_double, [C.c_void_p, C.c_int]),
            "orc_action_bead": (C.c_double, [C.c_void_p, C.c_int, C.c_int]),
            "orc_action_range": (C.c_double, [C.c_void_p, ip, ip]),
            "orc_action_list": (C.c_double, [C.c_void_p, ip, C.c_int]),
            "orc_update_bead": (None, [C.c_void_p, C.c_int, C.c_int, dp]),
            "orc_state_set_correction": (None, [C.c_void_p, C.c_double, C.c_double, C.c_int]),
            "orc_gradV2_slice": (C.c_double, [C.c_void_p, C.c_int]),
            "orc_gradVnn2": (C.c_double, [C.c_void_p, C.c_int, C.c_int]),
            "orc_bare_action_list": (None, [C.c_void_p, ip, C.c_int, dp]),
            "orc_action_correction": (C.c_double, [C.c_void_p, C.c_int, C.c_int]),
            "orc_energy": (None, [C.c_void_p, C.c_double, dp]),
            "orc_virial_terms": (None, [C.c_void_p, C.c_int, dp]),
            "orc_left_pack": (None, [C.c_int, C.c_int, dp, ip, ip, C.POINTER(C.c_uint32)]),
            "orc_virial_energy": (None, [C.c_void_p, C.c_int, C.c_double, dp]),
            "orc_ssf": (None, [C.c_void_p, dp, C.c_int, dp]),
            "orc_isf": (None, [C.c_void_p, dp, C.c_int, dp]),
            "orc_grid_dims": (None, [C.c_void_p, ip, ip]),
            "orc_grid_dump": (None, [C.c_void_p, ip, ip, ip, ip]),
            "orc_grid_nn": (None, [C.c_void_p, ip]),
            "orc_interaction_list": (C.c_int, [C.c_void_p, C.c_int, C.c_int, ip, dp]),
            "orc_qvectors": (C.c_int, [dp, C.c_char_p, C.c_char_p, dp, C.c_int]),
            "orc_gen_worldlines": (None, [C.c_uint32, C.c_double, C.c_int, C.c_int, C.c_double,
                                          C.c_double, C.c_int, dp]),
            "orc_moves_create": (C.c_void_p, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                              C.c_double]),
            "orc_moves_destroy": (None, [C.c_void_p]),
            "orc_move_bisection": (C.c_int, [C.c_void_p, C.c_void_p, dp]),
            "orc_move_staging": (C.c_int, [C.c_void_p, C.c_void_p, dp]),
            "orc_move_com": (C.c_int, [C.c_void_p, C.c_void_p, dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class MTRand:
    """The reference's MTRand (include/MersenneTwister.h)."""

    def __init__(self, seed):
        self.h = lib().orc_mt_create(seed)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mt_destroy(self.h)
            self.h = None

    def randInt(self, n=None):
        return lib().orc_mt_randint(self.h) if n is None else lib().orc_mt_randint_n(self.h, n)

    def rand(self):
        return lib().orc_mt_rand(self.h)

    def randExc(self):
        return lib().orc_mt_randexc(self.h)

    def randNorm(self, mean=0.0, sd=1.0):
        return lib().orc_mt_randnorm(self.h, mean, sd)

    def save(self):
        a = np.zeros(625, dtype=np.uint32)
        lib().orc_mt_save(self.h, _p(a, up))
        return a

    def load(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        lib().orc_mt_load(self.h, _p(a, up))


class Aziz:
    """AzizPotential + TabulatedPotential (src/potential.cpp:1741-1909)."""

    def __init__(self, year=1979, max_sep=10.0, rc=10.0, derivs=False, handle=None):
        self.owned = handle is None
        self.h = handle if handle is not None else lib().orc_aziz_create(year, max_sep, rc, int(derivs))

    def __del__(self):
        if getattr(self, "owned", False) and getattr(self, "h", None):
            lib().orc_aziz_destroy(self.h)
            self.h = None

    def valueV(self, r):
        return lib().orc_aziz_valueV(self.h, r)

    def valuedVdr(self, r):
        return lib().orc_aziz_valuedVdr(self.h, r)

    def valued2Vdr2(self, r):
        return lib().orc_aziz_valued2Vdr2(self.h, r)

    def V(self, sep):
        s = np.ascontiguousarray(sep, dtype=np.float64)
        return lib().orc_aziz_V(self.h, _p(s, dp))

    def gradV(self, sep):
        s = np.ascontiguousarray(sep, dtype=np.float64)
        o = np.zeros(3)
        lib().orc_aziz_gradV(self.h, _p(s, dp), _p(o, dp))
        return o

    def grad2V(self, sep):
        s = np.ascontiguousarray(sep, dtype=np.float64)
        return lib().orc_aziz_grad2V(self.h, _p(s, dp))

    @property
    def tailV(self):
        return lib().orc_aziz_tailV(self.h)

    @property
    def dr(self):
        return lib().orc_aziz_dr(self.h)

    @property
    def table_length(self):
        return lib().orc_aziz_table_length(self.h)

    def table(self, which=0):
        n = self.table_length
        ptr = lib().orc_aziz_table(self.h, which)
        if not ptr:
            return None
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    def branch_bins(self):
        o = np.zeros(3, dtype=np.int32)
        lib().orc_aziz_branch_bins(self.h, _p(o, ip))
        return o

    def params(self):
        o = np.zeros(9)
        lib().orc_aziz_params(self.h, _p(o, dp))
        return dict(zip(["epsilon", "rm", "A", "alpha", "beta", "D", "C6", "C8", "C10"], o))


class State:
    """Prism + Path + Worm + LookupTable + LocalAction in one object."""

    def __init__(self, side, M, nalloc, tau, lam, rc, year=1979):
        self.side, self.M, self.N, self.tau, self.lam, self.rc = side, M, nalloc, tau, lam, rc
        self.h = lib().orc_state_create(side, M, nalloc, tau, lam, rc, year)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_state_destroy(self.h)
            self.h = None

    @property
    def aziz(self):
        return Aziz(handle=lib().orc_state_aziz(self.h))

    def set_beads(self, beads):
        b = np.ascontiguousarray(beads, dtype=np.float64).reshape(self.M, self.N, 3)
        lib().orc_state_set_beads(self.h, _p(b, dp))

    def set_topology(self, bead_on, num_beads, nxt, prv, head=(-1, -1), tail=(-1, -1),
                     special1=(-1, -1), special2=(-1, -1)):
        on = np.ascontiguousarray(bead_on, dtype=np.uint8)
        nb = np.ascontiguousarray(num_beads, dtype=np.int32)
        nx = np.ascontiguousarray(nxt, dtype=np.int32)
        pv = np.ascontiguousarray(prv, dtype=np.int32)
        arrs = [np.array(x, dtype=np.int32) for x in (head, tail, special1, special2)]
        lib().orc_state_set_topology(self.h, _p(on, bp), _p(nb, ip), _p(nx, ip), _p(pv, ip),
                                     *[_p(a, ip) for a in arrs])

    def beads(self):
        b = np.zeros((self.M, self.N, 3))
        lib().orc_state_get_beads(self.h, _p(b, dp))
        return b

    def set_action(self, shift=1, vfactor=(1.0, 1.0)):
        lib().orc_state_set_action(self.h, shift, vfactor[0], vfactor[1])

    def set_correction(self, gradvfactor=(0.0, 0.0), period=1):
        """GSF / Li-Broughton gradVFactor[2] and estimator period (src/setup.cpp:1232-1256)."""
        lib().orc_state_set_correction(self.h, gradvfactor[0], gradvfactor[1], period)

    def gradV2_slice(self, slice_):
        return lib().orc_gradV2_slice(self.h, slice_)

    def gradVnn2(self, slice_, ptcl):
        return lib().orc_gradVnn2(self.h, slice_, ptcl)

    def bare_action_list(self, beads2):
        b = np.ascontiguousarray(beads2, dtype=np.int32).reshape(-1, 2)
        o = np.zeros(b.shape[0])
        lib().orc_bare_action_list(self.h, _p(b, ip), b.shape[0], _p(o, dp))
        return o

    def action_correction(self, slice_, ptcl):
        return lib().orc_action_correction(self.h, slice_, ptcl)

    def total_action(self, threads=None):
        """LocalAction::potentialAction(); threads: the slices' pair sums over that
        many host threads (the same value, primitive factors only)."""
        if threads is None:
            return lib().orc_total_action(self.h)
        return lib().orc_total_action_threads(self.h, threads)

    def V_slice(self, slice_, hist=False):
        o = np.zeros(2)
        h = np.zeros(50, dtype=np.int32)
        lib().orc_V_slice(self.h, slice_, _p(o, dp), _p(h, ip))
        return (o, h) if hist else o

    def Vnn(self, slice_, ptcl):
        return lib().orc_Vnn_bead(self.h, slice_, ptcl)

    def Vnn_slice(self, slice_):
        return lib().orc_Vnn_slice(self.h, slice_)

    def action_bead(self, slice_, ptcl):
        return lib().orc_action_bead(self.h, slice_, ptcl)

    def action_range(self, start, end):
        s = np.array(start, dtype=np.int32)
        e = np.array(end, dtype=np.int32)
        return lib().orc_action_range(self.h, _p(s, ip), _p(e, ip))

    def action_list(self, beads2):
        b = np.ascontiguousarray(beads2, dtype=np.int32).reshape(-1, 2)
        return lib().orc_action_list(self.h, _p(b, ip), b.shape[0])

    def update_bead(self, slice_, ptcl, pos):
        p = np.ascontiguousarray(pos, dtype=np.float64)
        lib().orc_update_bead(self.h, slice_, ptcl, _p(p, dp))

    def energy(self, mu=0.0):
        o = np.zeros(9)
        lib().orc_energy(self.h, mu, _p(o, dp))
        return dict(zip(ENERGY_COLUMNS, o))

    def virial_terms(self, window=5):
        """[M][6] = rDOTgradUterm1, rDOTgradUterm2, deltadotgradUterm1,
        deltadotgradUterm2, virialKinCorrection, secondderivPotentialActionTau."""
        o = np.zeros((self.M, 6))
        lib().orc_virial_terms(self.h, window, _p(o, dp))
        return o

    def virial_energy(self, window=5, mu=0.0):
        """VirialEnergyEstimator::accumulate, one call (src/estimator.cpp:1086-1251)."""
        o = np.zeros(19)
        lib().orc_virial_energy(self.h, window, mu, _p(o, dp))
        return dict(zip(VIRIAL_COLUMNS, o))

    def ssf(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
        o = np.zeros(q.shape[0])
        lib().orc_ssf(self.h, _p(q, dp), q.shape[0], _p(o, dp))
        return o

    def isf(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
        o = np.zeros((q.shape[0], self.M))
        lib().orc_isf(self.h, _p(q, dp), q.shape[0], _p(o, dp))
        return o

    def grid_dims(self):
        n = np.zeros(3, dtype=np.int32)
        cap = np.zeros(1, dtype=np.int32)
        lib().orc_grid_dims(self.h, _p(n, ip), _p(cap, ip))
        return n, int(cap[0])

    def grid_dump(self):
        n, cap = self.grid_dims()
        cells = int(np.prod(n))
        grid = np.zeros((self.M, self.N, 3), dtype=np.int32)
        lab = np.zeros((self.M, self.N), dtype=np.int32)
        nl = np.zeros((cells, self.M), dtype=np.int32)
        hs = np.zeros((cells, self.M, cap), dtype=np.int32)
        lib().orc_grid_dump(self.h, _p(grid, ip), _p(lab, ip), _p(nl, ip), _p(hs, ip))
        return dict(dims=n, cap=cap, grid=grid, bead_label=lab, num_labels=nl, hash=hs)

    def grid_nn(self):
        n, _ = self.grid_dims()
        o = np.zeros((int(np.prod(n)), 27, 3), dtype=np.int32)
        lib().orc_grid_nn(self.h, _p(o, ip))
        return o

    def interaction_list(self, slice_, ptcl):
        lst = np.zeros((self.N, 2), dtype=np.int32)
        sep = np.zeros((self.N, 3))
        n = lib().orc_interaction_list(self.h, slice_, ptcl, _p(lst, ip), _p(sep, dp))
        return lst[:n].copy(), sep[:n].copy()


class Moves:
    """Bisection / Staging / CenterOfMass attemptMove replays (src/move.cpp)."""

    def __init__(self, state, Mbar, b, var_updates=True, max_wind=1, com_delta=0.1):
        self.state = state
        self.h = lib().orc_moves_create(state.h, Mbar, b, int(var_updates), max_wind, com_delta)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_moves_destroy(self.h)
            self.h = None

    def _run(self, fn, rng):
        info = np.zeros(3)
        acc = fn(self.h, rng.h, _p(info, dp))
        return acc, info

    def bisection(self, rng):
        return self._run(lib().orc_move_bisection, rng)

    def staging(self, rng):
        return self._run(lib().orc_move_staging, rng)

    def com(self, rng):
        return self._run(lib().orc_move_com, rng)


def qvectors(side, kind, text, max_out=2_000_000):
    s = np.array([side, side, side] if np.isscalar(side) else side, dtype=np.float64)
    out = np.zeros((max_out, 3))
    n = lib().orc_qvectors(_p(s, dp), kind.encode(), text.encode(), _p(out, dp), max_out)
    if n < 0:
        raise ValueError("bad wavevector input")
    return out[:n].copy()


def gen_worldlines(seed, side, M, N, lam, tau, kind=0):
    b = np.zeros((M, N, 3))
    lib().orc_gen_worldlines(seed, side, M, N, lam, tau, kind, _p(b, dp))
    return b


def box_side(N, density):
    """Prism(density, N): side = pow(N/density, 1/3) (src/container.cpp:88)."""
    return (1.0 * N / density) ** (1.0 / 3.0)
