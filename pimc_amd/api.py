"""Python mirror of the C-ABI in include/pimc_gpu.h (ctypes over libpimc_gpu.so).

The classes mirror the reference's interfaces for the hot path:
  Context  ~ Path + LookupTable + LocalAction (src/action.cpp, src/path.cpp)
  Aziz     ~ AzizPotential as a batched PotentialBase (src/potential.cpp:1741-1909)
Methods keep the reference's names where one exists (potentialAction,
potential, Vnn-based action lists, updateBead) and raise PimcError on any
non-zero status, the way the reference aborts on a failed GPU_ASSERT.
Nothing here computes physics on the host: every result comes from the HIP
kernels, and importing the module fails loudly when the library is missing.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpimc_gpu.so")

V_FP32, V_FP64, V_TABLE = 0, 1, 2
# kernel names of pimc_kernel_time, in the bit order of pimc_set_timing_kinds
KERNEL_NAMES = ("pair_action", "delta_action", "rho", "ssf", "isf", "gradv2", "gradvnn", "virial", "chain_run")
ENERGY_COLUMNS = ("K", "V", "V_ext", "V_int", "E", "E_mu", "K/N", "V/N", "E/N")
VIRIAL_COLUMNS = ("K_op", "K_cv", "V_op", "V_cv", "E", "E_mu", "K_op/N", "K_cv/N", "V_op/N", "V_cv/N", "E/N",
                  "EEcv*Beta^2", "Ecv*Beta", "dEdB", "CvCov1", "CvCov2", "CvCov3", "E_th", "P")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_bp = C.POINTER(C.c_uint8)


class PimcError(RuntimeError):
    pass


class _Desc(C.Structure):
    _fields_ = [("side", C.c_double), ("num_slices", C.c_int), ("num_alloc", C.c_int),
                ("tau", C.c_double), ("lam", C.c_double), ("rc", C.c_double),
                ("aziz_year", C.c_int), ("v_mode", C.c_int)]


class _ServerDesc(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("streams", C.c_int), ("aziz_year", C.c_int), ("v_mode", C.c_int),
                ("max_side", C.c_double)]


class _LoadDesc(C.Structure):
    _fields_ = [("n_chains", C.c_int), ("rate", C.c_double), ("duration_s", C.c_double), ("seg_min", C.c_int),
                ("seg_max", C.c_int), ("p_com", C.c_double), ("com_sigma", C.c_double), ("lam", C.c_double),
                ("seed", C.c_uint32), ("threads", C.c_int), ("p_ssf", C.c_double), ("ssf_nmax", C.c_int)]


class _ChainDesc(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("Mbar", C.c_int), ("levels", C.c_int), ("var_updates", C.c_int),
                ("max_wind", C.c_int), ("com_delta", C.c_double)]


_vp = C.c_void_p
_u64p = C.POINTER(C.c_uint64)
_i64p = C.POINTER(C.c_int64)

SIGNATURES = {
    "pimc_ctx_create": (C.c_int, [C.POINTER(_Desc), C.POINTER(C.c_void_p)]),
    "pimc_ctx_destroy": (None, [C.c_void_p]),
    "pimc_last_error": (C.c_char_p, [C.c_void_p]),
    "pimc_set_v_mode": (C.c_int, [C.c_void_p, C.c_int]),
    "pimc_upload_path": (C.c_int, [C.c_void_p, _dp, _bp, _ip]),
    "pimc_set_worm": (C.c_int, [C.c_void_p, _ip, _ip, _ip, _ip]),
    "pimc_download_path": (C.c_int, [C.c_void_p, _dp]),
    "pimc_update_beads": (C.c_int, [C.c_void_p, _ip, _dp, C.c_int]),
    # the two per-move calls take raw addresses (c_void_p): no ctypes pointer objects per call
    "pimc_total_action": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pimc_potential_slices": (C.c_int, [C.c_void_p, _dp, _ip]),
    "pimc_energy": (C.c_int, [C.c_void_p, _ip, C.c_double, _dp]),
    "pimc_gradv_squared": (C.c_int, [C.c_void_p, _dp]),
    "pimc_path_version": (C.c_int, [C.c_void_p, _u64p]),
    "pimc_set_launch_shape": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "pimc_autotune": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "pimc_virial_terms": (C.c_int, [C.c_void_p, _ip, C.c_int, _dp]),
    "pimc_virial_energy": (C.c_int, [C.c_void_p, _ip, C.c_int, C.c_double, _dp]),
    "pimc_action_list": (C.c_int, [C.c_void_p, _ip, C.c_int, C.c_int, _dp, _dp, _dp]),
    "pimc_action_parts": (C.c_int, [C.c_void_p, _ip, C.c_int, C.c_int, _dp, _dp, _dp]),
    "pimc_action_range": (C.c_int, [C.c_void_p, _ip, _ip, _ip, C.c_int, _dp, _dp]),
    "pimc_delta_action": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                    C.c_int, C.c_void_p]),
    "pimc_ssf": (C.c_int, [C.c_void_p, _dp, C.c_int, _dp]),
    "pimc_ssf_noalloc": (C.c_int, [C.c_void_p, _dp, C.c_int, _dp]),
    "pimc_isf": (C.c_int, [C.c_void_p, _dp, C.c_int, _dp]),
    "pimc_cell_dims": (C.c_int, [C.c_void_p, _ip, _ip]),
    "pimc_cell_lists": (C.c_int, [C.c_void_p, _ip, _ip, _ip, _ip]),
    "pimc_interaction_list": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _ip, _dp, C.POINTER(C.c_int)]),
    "pimc_aziz_create": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_void_p)]),
    "pimc_aziz_destroy": (None, [C.c_void_p]),
    "pimc_aziz_V": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int]),
    "pimc_aziz_gradV": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int]),
    "pimc_aziz_grad2V": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int]),
    "pimc_aziz_tailV": (C.c_double, [C.c_void_p]),
    "pimc_aziz_table_length": (C.c_int, [C.c_void_p]),
    "pimc_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "pimc_set_timing_kinds": (C.c_int, [C.c_void_p, C.c_uint]),
    "pimc_kernel_time": (C.c_int, [C.c_void_p, C.c_char_p, _dp, C.POINTER(C.c_long)]),
    "pimc_device_synchronize": (C.c_int, []),
    "pimc_host_alloc": (C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    "pimc_host_free": (C.c_int, [C.c_void_p]),
    "pimc_copy_probe": (C.c_int, [C.c_size_t, C.c_int, _dp]),
    "pimc_gen_worldlines": (C.c_int, [C.c_uint32, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                      _dp]),
    "pimc_batch_create": (C.c_int, [C.POINTER(_Desc), C.c_int, C.POINTER(C.c_void_p)]),
    "pimc_batch_destroy": (None, [_vp]),
    "pimc_batch_last_error": (C.c_char_p, [_vp]),
    "pimc_batch_upload": (C.c_int, [_vp, C.c_int, _dp]),
    "pimc_batch_set_tiles": (C.c_int, [_vp, C.c_int, C.c_int]),
    "pimc_batch_use_graph": (C.c_int, [_vp, C.c_int]),
    "pimc_batch_total_action": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp]),
    "pimc_batch_info": (C.c_int, [_vp, _i64p, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp]),
    "pimc_batch_autotune": (C.c_int, [_vp, C.c_int, _dp]),
    "pimc_server_create": (C.c_int, [C.POINTER(_ServerDesc), C.POINTER(C.c_void_p)]),
    "pimc_server_destroy": (None, [_vp]),
    "pimc_server_last_error": (C.c_char_p, [_vp]),
    "pimc_server_add_chain": (C.c_int, [_vp, C.POINTER(_Desc), _dp, C.POINTER(C.c_int)]),
    "pimc_server_chain_state": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int)]),
    "pimc_server_remove_chain": (C.c_int, [_vp, C.c_int]),
    "pimc_server_download": (C.c_int, [_vp, C.c_int, _dp]),
    "pimc_server_submit": (C.c_int, [_vp, C.c_int, _ip, _dp, C.c_int, C.c_int, _dp, C.c_int, C.c_double, _u64p]),
    "pimc_server_poll": (C.c_int, [_vp, C.c_uint64, C.POINTER(C.c_int)]),
    "pimc_server_wait": (C.c_int, [_vp, C.c_uint64, _dp]),
    "pimc_server_submit_ssf": (C.c_int, [_vp, C.c_int, _dp, C.c_int, _u64p]),
    "pimc_server_wait_ssf": (C.c_int, [_vp, C.c_uint64, _dp, _dp]),
    "pimc_server_stats": (C.c_int, [_vp, _i64p, _i64p, _i64p, C.POINTER(C.c_size_t)]),
    "pimc_server_dispatch_stats": (C.c_int, [_vp, _i64p, _i64p]),
    "pimc_server_run_synthetic": (C.c_int, [_vp, C.POINTER(_LoadDesc), _dp]),
    "pimc_set_action_factors": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int]),
    "pimc_chain_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "pimc_chain_destroy": (None, [C.c_void_p]),
    "pimc_chain_last_error": (C.c_char_p, [C.c_void_p]),
    "pimc_chain_move": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), _dp]),
    "pimc_chain_positions": (C.c_int, [C.c_void_p, _dp]),
    "pimc_chain_run_device": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _dp, _ip]),
    "pimc_chain_run_device_many": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int, _ip, _ip, _dp, _ip]),
    "pimc_state_last_error": (C.c_char_p, []),
    "pimc_state_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "pimc_state_parse": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "pimc_state_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "pimc_state_destroy": (None, [C.c_void_p]),
    "pimc_state_dims": (C.c_int, [C.c_void_p, _ip, _ip, _ip]),
    "pimc_state_get": (C.c_int, [C.c_void_p, _dp, _ip, _ip, C.POINTER(C.c_uint32)]),
    "pimc_state_set": (C.c_int, [C.c_void_p, _dp, _ip, _ip, C.POINTER(C.c_uint32)]),
    "pimc_state_counters": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.c_int, _ip, C.POINTER(C.c_int64),
                                      C.c_int]),
    "pimc_state_random": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), _ip, C.POINTER(C.c_uint64)]),
    "pimc_state_left_pack": (C.c_int, [C.c_void_p]),
    "pimc_state_format": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "pimc_state_write": (C.c_int, [C.c_void_p, C.c_char_p]),
    "pimc_state_to_path": (C.c_int, [C.c_void_p, C.c_int, C.c_double, _dp, _bp, _ip, _ip, _ip]),
}

_lib = None


def lib():
    """Load libpimc_gpu.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PimcError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def _ptr(a):
    """Address of a C-contiguous array's data: ctypes.addressof over the buffer
    (about 1 us), or the ndarray.ctypes route for read-only arrays."""
    try:
        return C.addressof(C.c_char.from_buffer(a))
    except (TypeError, ValueError, BufferError):
        return a.ctypes.data


def _i32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a if shape is None else a.reshape(shape)


class Context:
    """One chain's worldlines resident in HBM (include/pimc_gpu.h, pimc_ctx)."""

    def __init__(self, side, num_slices, num_alloc, tau, lam, rc, year=1979, v_mode=V_FP32):
        d = _Desc(side, num_slices, num_alloc, tau, lam, rc, year, v_mode)
        h = C.c_void_p()
        rc_ = lib().pimc_ctx_create(C.byref(d), C.byref(h))
        if rc_ != 0:
            raise PimcError(f"pimc_ctx_create: {lib().pimc_last_error(None).decode()} ({rc_})")
        self.h = h
        self.M, self.N, self.side = num_slices, num_alloc, side
        self._out = np.zeros(4)         # reused result buffer of the per-move calls
        self._out_ptr = self._out.ctypes.data
        self._fn_dv = lib().pimc_delta_action
        self._fn_ta = lib().pimc_total_action

    def close(self):
        if getattr(self, "h", None):
            lib().pimc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise PimcError(f"{what}: {lib().pimc_last_error(self.h).decode()} ({rc})")

    def set_action_factors(self, vfactor=(1.0, 1.0), gradvfactor=(0.0, 0.0), period=1):
        """LocalAction VFactor, gradVFactor (GSF / Li-Broughton correction) and estimator period."""
        vf, gf = _f64(vfactor), _f64(gradvfactor)
        self._check(lib().pimc_set_action_factors(self.h, _p(vf, _dp), _p(gf, _dp), period), "set_action_factors")

    def set_launch_shape(self, pairs_per_lane=0, grid_per_cu=0, stage_in_lds=-1):
        self._check(lib().pimc_set_launch_shape(self.h, pairs_per_lane, grid_per_cu, stage_in_lds),
                    "set_launch_shape")

    def autotune(self, reps=5):
        """Fastest launch shape of the all-pairs kernel: dict(pairs_per_lane, grid_per_cu, stage_in_lds, ms)."""
        o = np.zeros(4)
        self._check(lib().pimc_autotune(self.h, reps, _p(o, _dp)), "autotune")
        return dict(pairs_per_lane=int(o[0]), grid_per_cu=int(o[1]), stage_in_lds=int(o[2]), ms=o[3])

    def set_v_mode(self, mode):
        self._check(lib().pimc_set_v_mode(self.h, mode), "set_v_mode")

    def upload_path(self, beads, bead_on=None, num_beads=None):
        b = _f64(beads, (self.M, self.N, 3))
        on = None if bead_on is None else np.ascontiguousarray(bead_on, dtype=np.uint8).reshape(self.M, self.N)
        nb = None if num_beads is None else _i32(num_beads, (self.M,))
        self._check(lib().pimc_upload_path(self.h, _p(b, _dp), _p(on, _bp), _p(nb, _ip)), "upload_path")

    def load_state(self, state):
        """loadState into this context (src/pimc.cpp:1108-1258): the path from a
        StateFile, uploaded with its active-bead mask and counts, the worm reset.
        Returns the path dict (next / prev links for energy and virial calls)."""
        p = state.to_path(self.M, self.side)
        if p["beads"].shape[1] != self.N:
            raise PimcError(f"load_state: the state has {p['beads'].shape[1]} worldlines, the context {self.N}")
        self.upload_path(p["beads"], p["bead_on"], p["num_beads"])
        self.set_worm()
        return p

    def set_worm(self, head=(-1, -1), tail=(-1, -1), special1=(-1, -1), special2=(-1, -1)):
        arrs = [_i32(x, (2,)) for x in (head, tail, special1, special2)]
        self._check(lib().pimc_set_worm(self.h, *[_p(a, _ip) for a in arrs]), "set_worm")

    def download_path(self):
        b = np.zeros((self.M, self.N, 3))
        self._check(lib().pimc_download_path(self.h, _p(b, _dp)), "download_path")
        return b

    def update_beads(self, slice_ptcl, pos):
        sp = _i32(slice_ptcl, (-1, 2))
        p = _f64(pos, (-1, 3))
        self._check(lib().pimc_update_beads(self.h, _p(sp, _ip), _p(p, _dp), sp.shape[0]), "update_beads")

    def potentialAction(self, shift=1, vfactor=None, with_slices=False):
        """vfactor None: the context's VFactor (set_action_factors; {1, 1} by default)."""
        if vfactor is None and not with_slices:
            # the per-step call: one ctypes call into the reused result buffer
            rc = self._fn_ta(self.h, shift, None, self._out_ptr, None)
            if rc != 0:
                self._check(rc, "total_action")
            return float(self._out[0])
        vf = None if vfactor is None else _f64(vfactor)
        vs = np.zeros((self.M, 2)) if with_slices else None
        self._check(lib().pimc_total_action(self.h, shift, None if vf is None else _ptr(vf), self._out_ptr,
                                            None if vs is None else _ptr(vs)), "total_action")
        a = float(self._out[0])
        return (a, vs) if with_slices else a

    def potential_slices(self, sep_hist=False):
        vs = np.zeros((self.M, 2))
        h = np.zeros(50, dtype=np.int32) if sep_hist else None
        self._check(lib().pimc_potential_slices(self.h, _p(vs, _dp), _p(h, _ip)), "potential_slices")
        return (vs, h) if sep_hist else vs

    def energy(self, next_links=None, mu=0.0):
        """EnergyEstimator::accumulate, the reference's nine columns (src/estimator.cpp:922-1016)."""
        o = np.zeros(9)
        nx = None if next_links is None else _i32(next_links, (self.M, self.N, 2))
        self._check(lib().pimc_energy(self.h, _p(nx, _ip), mu, _p(o, _dp)), "energy")
        return dict(zip(ENERGY_COLUMNS, o))

    def gradVSquared(self):
        """LocalAction::gradVSquared(slice) for every slice [M] (src/action.cpp:1188-1223)."""
        o = np.zeros(self.M)
        self._check(lib().pimc_gradv_squared(self.h, _p(o, _dp)), "gradv_squared")
        return o

    def path_version(self):
        v = C.c_uint64()
        self._check(lib().pimc_path_version(self.h, C.byref(v)), "path_version")
        return v.value

    def virial_terms(self, window=5, next_links=None):
        """[M][6]: rDOTgradUterm1, rDOTgradUterm2, deltadotgradUterm1,
        deltadotgradUterm2, virialKinCorrection, secondderivPotentialActionTau
        (src/action.cpp:1446-1786, :776-787)."""
        o = np.zeros((self.M, 6))
        nx = None if next_links is None else _i32(next_links, (self.M, self.N, 2))
        self._check(lib().pimc_virial_terms(self.h, _p(nx, _ip), window, _p(o, _dp)), "virial_terms")
        return o

    def virial_energy(self, window=5, mu=0.0, next_links=None):
        """VirialEnergyEstimator::accumulate, one call (src/estimator.cpp:1086-1251)."""
        o = np.zeros(19)
        nx = None if next_links is None else _i32(next_links, (self.M, self.N, 2))
        self._check(lib().pimc_virial_energy(self.h, _p(nx, _ip), window, mu, _p(o, _dp)), "virial_energy")
        return dict(zip(VIRIAL_COLUMNS, o))

    def action_list(self, slice_ptcl, shift=1, vfactor=None, per_bead=False):
        sp = _i32(slice_ptcl, (-1, 2))
        vf = None if vfactor is None else _f64(vfactor)
        a = np.zeros(1)
        pb = np.zeros(sp.shape[0]) if per_bead else None
        self._check(lib().pimc_action_list(self.h, _p(sp, _ip), sp.shape[0], shift, _p(vf, _dp), _p(a, _dp),
                                           _p(pb, _dp)), "action_list")
        return (a[0], pb) if per_bead else a[0]

    def action_parts(self, slice_ptcl, shift=1, vfactor=None):
        """(bare [n], correction [n]) per bead: barePotentialAction and
        potentialActionCorrection (src/action.cpp:647-716)."""
        sp = _i32(slice_ptcl, (-1, 2))
        vf = None if vfactor is None else _f64(vfactor)
        bare, cor = np.zeros(sp.shape[0]), np.zeros(sp.shape[0])
        self._check(lib().pimc_action_parts(self.h, _p(sp, _ip), sp.shape[0], shift, _p(vf, _dp), _p(bare, _dp),
                                            _p(cor, _dp)), "action_parts")
        return bare, cor

    def barePotentialAction(self, slice_ptcl, shift=1):
        """LocalAction::barePotentialAction(beads, values) (src/action.cpp:647-689)."""
        return self.action_parts(slice_ptcl, shift)[0]

    def potentialActionCorrection(self, slice_ptcl, shift=1):
        """LocalAction::potentialActionCorrection summed over the listed beads
        (src/action.cpp:696-738; a start..end range is its bead list)."""
        return float(sum(self.action_parts(slice_ptcl, shift)[1].tolist()))   # in bead order

    def action_range(self, start, end, next_links=None, shift=1, vfactor=None):
        s, e = _i32(start), _i32(end)
        nx = None if next_links is None else _i32(next_links, (self.M, self.N, 2))
        vf = None if vfactor is None else _f64(vfactor)
        a = np.zeros(1)
        self._check(lib().pimc_action_range(self.h, _p(s, _ip), _p(e, _ip), _p(nx, _ip), shift, _p(vf, _dp),
                                            _p(a, _dp)), "action_range")
        return a[0]

    def delta_action(self, slice_ptcl, new_pos, old_pos=None, shift=1, vfactor=None, commit=False):
        sp = _i32(slice_ptcl, (-1, 2))
        npos = _f64(new_pos, (-1, 3))
        opos = None if old_pos is None else _f64(old_pos, (-1, 3))
        if npos.shape[0] != sp.shape[0] or (opos is not None and opos.shape[0] != sp.shape[0]):
            raise PimcError("delta_action: bead list and positions differ in length")
        vf = None if vfactor is None else _f64(vfactor)
        self._check(lib().pimc_delta_action(self.h, _ptr(sp), None if opos is None else _ptr(opos), _ptr(npos),
                                            sp.shape[0], shift, None if vf is None else _ptr(vf),
                                            int(bool(commit)), self._out_ptr), "delta_action")
        return self._out[:3].copy()

    def prepare_delta(self, slice_ptcl, new_pos):
        """A Delta V request checked and converted once, for repeated
        delta_action_prepared calls (the arrays are kept alive with it)."""
        sp = _i32(slice_ptcl, (-1, 2))
        npos = _f64(new_pos, (-1, 3))
        if npos.shape[0] != sp.shape[0]:
            raise PimcError("delta_action: bead list and positions differ in length")
        return (sp, npos, _ptr(sp), _ptr(npos), sp.shape[0])

    def delta_action_prepared(self, req, commit=False):
        """delta_action for a prepare_delta request at shift 1 with the context's
        factors: {A_old, A_new, dU} in the context's reused result buffer (valid
        until the next call)."""
        rc = self._fn_dv(self.h, req[2], None, req[3], req[4], 1, None, int(bool(commit)), self._out_ptr)
        if rc != 0:
            self._check(rc, "delta_action")
        return self._out

    def ssf(self, q):
        q = _f64(q, (-1, 3))
        o = np.zeros(q.shape[0])
        self._check(lib().pimc_ssf(self.h, _p(q, _dp), q.shape[0], _p(o, _dp)), "ssf")
        return o

    def ssf_noalloc(self, q):
        """S(q) through pimc_ssf_noalloc: no device or host allocation in the library."""
        q = _f64(q, (-1, 3))
        o = np.zeros(q.shape[0])
        self._check(lib().pimc_ssf_noalloc(self.h, _p(q, _dp), q.shape[0], _p(o, _dp)), "ssf_noalloc")
        return o

    def isf(self, q, out=None):
        """F(q, tau) [nq][M]; out: a C-contiguous float64 array of that shape to
        write into (pinned_empty gives one in page-locked memory)."""
        q = _f64(q, (-1, 3))
        if out is None:
            o = np.empty((q.shape[0], self.M))   # every row is written
        else:
            o = out
            if o.dtype != np.float64 or o.shape != (q.shape[0], self.M) or not o.flags.c_contiguous:
                raise PimcError(f"isf: out must be a C-contiguous float64 array of shape ({q.shape[0]}, {self.M})")
        self._check(lib().pimc_isf(self.h, _p(q, _dp), q.shape[0], _p(o, _dp)), "isf")
        return o

    def cell_dims(self):
        d = np.zeros(3, dtype=np.int32)
        cap = np.zeros(1, dtype=np.int32)
        self._check(lib().pimc_cell_dims(self.h, _p(d, _ip), _p(cap, _ip)), "cell_dims")
        return d, int(cap[0])

    def cell_lists(self):
        d, cap = self.cell_dims()
        cells = int(np.prod(d))
        grid = np.zeros((self.M, self.N, 3), dtype=np.int32)
        lab = np.zeros((self.M, self.N), dtype=np.int32)
        nl = np.zeros((cells, self.M), dtype=np.int32)
        hs = np.zeros((cells, self.M, cap), dtype=np.int32)
        self._check(lib().pimc_cell_lists(self.h, _p(grid, _ip), _p(lab, _ip), _p(nl, _ip), _p(hs, _ip)),
                    "cell_lists")
        return dict(dims=d, cap=cap, grid=grid, bead_label=lab, num_labels=nl, hash=hs)

    def interaction_list(self, slc, ptcl):
        lst = np.zeros((self.N, 2), dtype=np.int32)
        sep = np.zeros((self.N, 3))
        cnt = C.c_int()
        self._check(lib().pimc_interaction_list(self.h, slc, ptcl, _p(lst, _ip), _p(sep, _dp), C.byref(cnt)),
                    "interaction_list")
        return lst[:cnt.value].copy(), sep[:cnt.value].copy()

    def set_timing(self, on=True, kinds=None):
        """Event timing of the hot kernels; kinds: names to time (None: all)."""
        if kinds is None or not on:
            self._check(lib().pimc_set_timing(self.h, int(on)), "set_timing")
            return
        mask = 0
        for k in kinds:
            mask |= 1 << KERNEL_NAMES.index(k)
        self._check(lib().pimc_set_timing_kinds(self.h, mask), "set_timing_kinds")

    def kernel_time(self, name):
        ms = C.c_double()
        cnt = C.c_long()
        self._check(lib().pimc_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(cnt)), "kernel_time")
        return ms.value, cnt.value


class Aziz:
    """Batched PotentialBase for the Aziz pair potential (include/pimc_gpu.h)."""

    def __init__(self, year, max_sep, rc, v_mode=V_FP32):
        h = C.c_void_p()
        rc_ = lib().pimc_aziz_create(year, max_sep, rc, v_mode, C.byref(h))
        if rc_ != 0:
            raise PimcError(f"pimc_aziz_create failed ({rc_})")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().pimc_aziz_destroy(self.h)
            self.h = None

    def _run(self, fn, sep, width):
        s = _f64(sep, (-1, 3))
        o = np.zeros((s.shape[0], width)) if width > 1 else np.zeros(s.shape[0])
        rc_ = fn(self.h, _p(s, _dp), _p(o, _dp), s.shape[0])
        if rc_ != 0:
            raise PimcError(f"aziz batch failed ({rc_})")
        return o

    def V(self, sep):
        return self._run(lib().pimc_aziz_V, sep, 1)

    def gradV(self, sep):
        return self._run(lib().pimc_aziz_gradV, sep, 3)

    def grad2V(self, sep):
        return self._run(lib().pimc_aziz_grad2V, sep, 1)

    @property
    def tailV(self):
        return lib().pimc_aziz_tailV(self.h)

    @property
    def table_length(self):
        return lib().pimc_aziz_table_length(self.h)


class Batch:
    """Batched form: potentialAction() of many walkers in one persistent launch."""

    def __init__(self, walkers, year=1979, v_mode=V_FP32):
        """walkers: list of dicts with side, M, N, tau, lam, rc."""
        arr = (_Desc * len(walkers))()
        for k, w in enumerate(walkers):
            arr[k] = _Desc(w["side"], w["M"], w["N"], w["tau"], w["lam"], w["rc"], year, v_mode)
        h = C.c_void_p()
        rc_ = lib().pimc_batch_create(arr, len(walkers), C.byref(h))
        if rc_ != 0:
            raise PimcError(f"pimc_batch_create failed ({rc_})")
        self.h = h
        self.walkers = walkers

    def __del__(self):
        if getattr(self, "h", None):
            lib().pimc_batch_destroy(self.h)
            self.h = None

    def _check(self, rc, what):
        if rc != 0:
            raise PimcError(f"{what}: {lib().pimc_batch_last_error(self.h).decode()} ({rc})")

    def upload(self, w, beads):
        b = _f64(beads, (self.walkers[w]["M"], self.walkers[w]["N"], 3))
        self._check(lib().pimc_batch_upload(self.h, w, _p(b, _dp)), "batch_upload")

    def set_tiles(self, pairs_per_lane, grid=0):
        self._check(lib().pimc_batch_set_tiles(self.h, pairs_per_lane, grid), "batch_set_tiles")

    def use_graph(self, on=True):
        self._check(lib().pimc_batch_use_graph(self.h, int(on)), "batch_use_graph")

    def potentialAction(self, shift=1, vfactor=(1.0, 1.0), with_slices=False):
        vf = _f64(vfactor)
        a = np.zeros(len(self.walkers))
        ns = sum(w["M"] for w in self.walkers)
        vs = np.zeros(ns) if with_slices else None
        self._check(lib().pimc_batch_total_action(self.h, shift, _p(vf, _dp), _p(a, _dp), _p(vs, _dp)),
                    "batch_total_action")
        return (a, vs) if with_slices else a

    def info(self):
        tp = C.c_int64()
        ni, ng = C.c_int(), C.c_int()
        ms = np.zeros(1)
        self._check(lib().pimc_batch_info(self.h, C.byref(tp), C.byref(ni), C.byref(ng), _p(ms, _dp)), "batch_info")
        return dict(total_pairs=tp.value, n_items=ni.value, n_groups=ng.value, last_ms=float(ms[0]))

    def autotune(self, reps=3):
        o = np.zeros(4)
        self._check(lib().pimc_batch_autotune(self.h, reps, _p(o, _dp)), "batch_autotune")
        return dict(pairs_per_lane=int(o[0]), grid=int(o[1]), ms=float(o[2]), candidates=int(o[3]))


class Server:
    """Serving form: Delta V requests of many resident chains, continuous batching."""

    def __init__(self, hbm_budget, streams=2, year=1979, v_mode=V_FP32, max_side=60.0):
        d = _ServerDesc(int(hbm_budget), streams, year, v_mode, max_side)
        h = C.c_void_p()
        rc_ = lib().pimc_server_create(C.byref(d), C.byref(h))
        if rc_ != 0:
            raise PimcError(f"pimc_server_create failed ({rc_})")
        self.h = h
        self.year, self.v_mode = year, v_mode
        self.shapes = {}
        self._nq = {}

    def close(self):
        if getattr(self, "h", None):
            lib().pimc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise PimcError(f"{what}: {lib().pimc_server_last_error(self.h).decode()} ({rc})")

    def add_chain(self, side, M, N, tau, lam, rc, beads):
        d = _Desc(side, M, N, tau, lam, rc, self.year, self.v_mode)
        b = _f64(beads, (M, N, 3))
        cid = C.c_int()
        self._check(lib().pimc_server_add_chain(self.h, C.byref(d), _p(b, _dp), C.byref(cid)), "add_chain")
        self.shapes[cid.value] = (M, N)
        return cid.value

    def chain_state(self, chain):
        st = C.c_int()
        self._check(lib().pimc_server_chain_state(self.h, chain, C.byref(st)), "chain_state")
        return st.value

    def remove_chain(self, chain):
        self._check(lib().pimc_server_remove_chain(self.h, chain), "remove_chain")

    def download(self, chain):
        M, N = self.shapes[chain]
        b = np.zeros((M, N, 3))
        self._check(lib().pimc_server_download(self.h, chain, _p(b, _dp)), "download")
        return b

    def submit(self, chain, slice_ptcl, new_pos, shift=1, vfactor=(1.0, 1.0), commit_mode=0, u=0.0):
        sp = _i32(slice_ptcl, (-1, 2))
        npos = _f64(new_pos, (-1, 3))
        vf = _f64(vfactor)
        rid = C.c_uint64()
        self._check(lib().pimc_server_submit(self.h, chain, _p(sp, _ip), _p(npos, _dp), sp.shape[0], shift,
                                             _p(vf, _dp), commit_mode, u, C.byref(rid)), "submit")
        return rid.value

    def poll(self, rid):
        d = C.c_int()
        self._check(lib().pimc_server_poll(self.h, rid, C.byref(d)), "poll")
        return bool(d.value)

    def wait(self, rid):
        o = np.zeros(5)
        self._check(lib().pimc_server_wait(self.h, rid, _p(o, _dp)), "wait")
        return dict(A_old=o[0], A_new=o[1], dU=o[2], accepted=bool(o[3] > 0.5), latency_us=o[4], raw=o)

    def submit_ssf(self, chain, q):
        """S(q) of a resident chain (lattice q of its box), queued behind its earlier requests."""
        q = _f64(q, (-1, 3))
        rid = C.c_uint64()
        self._check(lib().pimc_server_submit_ssf(self.h, chain, _p(q, _dp), q.shape[0], C.byref(rid)),
                    "submit_ssf")
        self._nq[rid.value] = q.shape[0]
        return rid.value

    def wait_ssf(self, rid):
        """(S(q) [nq], latency_us) of an S(q) request."""
        o = np.zeros(self._nq.pop(rid))
        lat = np.zeros(1)
        self._check(lib().pimc_server_wait_ssf(self.h, rid, _p(o, _dp), _p(lat, _dp)), "wait_ssf")
        return o, float(lat[0])

    def stats(self):
        b, r, p = C.c_int64(), C.c_int64(), C.c_int64()
        f = C.c_size_t()
        self._check(lib().pimc_server_stats(self.h, C.byref(b), C.byref(r), C.byref(p), C.byref(f)), "stats")
        return dict(batches=b.value, requests=r.value, pair_evals=p.value, free_bytes=f.value)

    def dispatch_stats(self):
        """Delta V batches sent as AQL packets and through hipLaunchKernel."""
        pk, hp = C.c_int64(), C.c_int64()
        self._check(lib().pimc_server_dispatch_stats(self.h, C.byref(pk), C.byref(hp)), "dispatch_stats")
        return dict(packet_batches=pk.value, hip_batches=hp.value)

    def run_synthetic(self, n_chains, rate, duration_s, seg=(4, 64), p_com=0.1, com_sigma=0.1, lam=6.0596,
                      seed=139853, threads=1, p_ssf=0.0, ssf_nmax=4):
        d = _LoadDesc(n_chains, rate, duration_s, seg[0], seg[1], p_com, com_sigma, lam, seed, threads, p_ssf,
                      ssf_nmax)
        o = np.zeros(16)
        self._check(lib().pimc_server_run_synthetic(self.h, C.byref(d), _p(o, _dp)), "run_synthetic")
        keys = ["completed", "req_per_s", "pair_evals_per_s", "p50_us", "p99_us", "mean_us", "accept_frac",
                "batches", "req_per_batch", "seconds", "service_p50_us", "service_p99_us",
                "ssf_completed", "ssf_p50_us", "ssf_p99_us", "ssf_phasors_per_s"]
        return dict(zip(keys, o))


class Chain:
    """One Markov chain's Staging / Bisection / CenterOfMass moves on a Context
    (include/pimc_gpu.h, pimc_chain): MTRand(seed) proposals and Metropolis tests
    on the host, Delta V on the device."""

    KINDS = {"staging": 0, "bisection": 1, "com": 2}

    def __init__(self, ctx, seed, Mbar=8, levels=3, var_updates=True, max_wind=1, com_delta=0.1):
        d = _ChainDesc(seed, Mbar, levels, int(bool(var_updates)), max_wind, com_delta)
        h = C.c_void_p()
        rc = lib().pimc_chain_create(ctx.h, C.byref(d), C.byref(h))
        if rc != 0:
            raise PimcError(f"pimc_chain_create failed ({rc}): {lib().pimc_last_error(ctx.h).decode()}")
        self.h, self.ctx = h, ctx
        self._acc = C.c_int()
        self._info = np.zeros(3)

    def close(self):
        if getattr(self, "h", None):
            lib().pimc_chain_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def move(self, kind):
        """kind: 'staging' | 'bisection' | 'com'.  Returns (accepted, info3)."""
        rc = lib().pimc_chain_move(self.h, self.KINDS[kind], C.byref(self._acc), _p(self._info, _dp))
        if rc != 0:
            raise PimcError(f"chain move ({rc}): {lib().pimc_chain_last_error(self.h).decode()}")
        return bool(self._acc.value), self._info.copy()

    def run_device(self, kinds):
        """The moves of `kinds` (names or 0/1/2) in one device launch.  Returns
        (accepted [n], info [n][3], near_threshold)."""
        k = np.array([self.KINDS[x] if isinstance(x, str) else int(x) for x in kinds], dtype=np.int32)
        acc = np.zeros(k.shape[0], np.int32)
        info = np.zeros((k.shape[0], 3))
        near = C.c_int()
        rc = lib().pimc_chain_run_device(self.h, k.shape[0], _p(k, _ip), _p(acc, _ip), _p(info, _dp),
                                         C.byref(near))
        if rc != 0:
            raise PimcError(f"chain run_device ({rc}): {lib().pimc_chain_last_error(self.h).decode()}")
        return acc.astype(bool), info, near.value

    def positions(self):
        b = np.zeros((self.ctx.M, self.ctx.N, 3))
        rc = lib().pimc_chain_positions(self.h, _p(b, _dp))
        if rc != 0:
            raise PimcError(f"chain positions ({rc})")
        return b


class StateFile:
    """The reference's text state file (include/pimc_gpu.h, pimc_state):
    PathIntegralMonteCarlo::saveState / loadState, src/pimc.cpp:926-986,
    :1108-1258.  Host code in the library."""

    def __init__(self, handle):
        self.h = handle

    @staticmethod
    def _make(rc, h, what):
        if rc != 0:
            raise PimcError(f"{what}: {lib().pimc_state_last_error().decode()} ({rc})")
        return StateFile(h)

    @classmethod
    def read(cls, path):
        h = C.c_void_p()
        return cls._make(lib().pimc_state_read(os.fsencode(path), C.byref(h)), h, "state read")

    @classmethod
    def parse(cls, text):
        b = text.encode() if isinstance(text, str) else bytes(text)
        h = C.c_void_p()
        return cls._make(lib().pimc_state_parse(b, len(b), C.byref(h)), h, "state parse")

    @classmethod
    def create(cls, num_slices, num_alloc):
        h = C.c_void_p()
        return cls._make(lib().pimc_state_create(num_slices, num_alloc, C.byref(h)), h, "state create")

    def close(self):
        if getattr(self, "h", None):
            lib().pimc_state_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise PimcError(f"{what}: {lib().pimc_state_last_error().decode()} ({rc})")

    def dims(self):
        m, n, w = C.c_int(), C.c_int(), C.c_int()
        self._check(lib().pimc_state_dims(self.h, C.byref(m), C.byref(n), C.byref(w)), "state dims")
        return m.value, n.value, w.value

    def arrays(self):
        M, N, _ = self.dims()
        b, nx, pv = np.zeros((M, N, 3)), np.zeros((M, N, 2), np.int32), np.zeros((M, N, 2), np.int32)
        w = np.zeros((M, N), np.uint32)
        self._check(lib().pimc_state_get(self.h, _p(b, _dp), _p(nx, _ip), _p(pv, _ip),
                                         w.ctypes.data_as(C.POINTER(C.c_uint32))), "state get")
        return {"beads": b, "next": nx, "prev": pv, "worm": w}

    def set(self, beads=None, next=None, prev=None, worm=None):
        M, N, _ = self.dims()
        b = None if beads is None else _f64(beads, (M, N, 3))
        nx = None if next is None else _i32(next, (M, N, 2))
        pv = None if prev is None else _i32(prev, (M, N, 2))
        w = None if worm is None else np.ascontiguousarray(worm, dtype=np.uint32).reshape(M, N)
        self._check(lib().pimc_state_set(self.h, _p(b, _dp), _p(nx, _ip), _p(pv, _ip),
                                         None if w is None else w.ctypes.data_as(C.POINTER(C.c_uint32))),
                    "state set")

    def counters(self):
        n = C.c_int()
        self._check(lib().pimc_state_counters(self.h, None, 0, C.byref(n), None, -1), "state counters")
        a = np.zeros((n.value, 2), np.int64)
        self._check(lib().pimc_state_counters(self.h, a.ctypes.data_as(C.POINTER(C.c_int64)), n.value, None,
                                              None, -1), "state counters")
        return a

    def set_counters(self, pairs):
        a = np.ascontiguousarray(pairs, dtype=np.int64).reshape(-1, 2)
        self._check(lib().pimc_state_counters(self.h, None, 0, None, a.ctypes.data_as(C.POINTER(C.c_int64)),
                                              a.shape[0]), "state counters")

    def random(self):
        w, have = np.zeros(625, np.uint64), C.c_int()
        self._check(lib().pimc_state_random(self.h, w.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(have),
                                            None), "state random")
        return w if have.value else None

    def set_random(self, words):
        w = np.ascontiguousarray(words, dtype=np.uint64).reshape(625)
        self._check(lib().pimc_state_random(self.h, None, None, w.ctypes.data_as(C.POINTER(C.c_uint64))),
                    "state random")

    def left_pack(self):
        self._check(lib().pimc_state_left_pack(self.h), "leftPack")

    def text(self):
        n = C.c_size_t()
        self._check(lib().pimc_state_format(self.h, None, 0, C.byref(n)), "state format")
        buf = C.create_string_buffer(n.value + 1)
        self._check(lib().pimc_state_format(self.h, buf, n.value + 1, None), "state format")
        return buf.raw[:n.value].decode()

    def write(self, path):
        self._check(lib().pimc_state_write(self.h, os.fsencode(path)), "state write")

    def to_path(self, num_slices, side):
        """loadState's set-up for a run of num_slices slices in a cube of side `side`."""
        _, _, W = self.dims()
        M = num_slices
        b, on = np.zeros((M, W, 3)), np.zeros((M, W), np.uint8)
        nb, nx, pv = np.zeros(M, np.int32), np.zeros((M, W, 2), np.int32), np.zeros((M, W, 2), np.int32)
        self._check(lib().pimc_state_to_path(self.h, M, side, _p(b, _dp), _p(on, _bp), _p(nb, _ip), _p(nx, _ip),
                                             _p(pv, _ip)), "state to path")
        return {"beads": b, "bead_on": on, "num_beads": nb, "next": nx, "prev": pv}


def run_chains_device(chains, kinds):
    """pimc_chain_run_device_many: every chain's moves of `kinds` in one launch.
    Returns (accepted [n][k], info [n][k][3], near_threshold [n])."""
    k = np.array([Chain.KINDS[x] if isinstance(x, str) else int(x) for x in kinds], dtype=np.int32)
    n = len(chains)
    hs = (C.c_void_p * n)(*[c.h.value for c in chains])
    acc = np.zeros((n, k.shape[0]), np.int32)
    info = np.zeros((n, k.shape[0], 3))
    near = np.zeros(n, np.int32)
    rc = lib().pimc_chain_run_device_many(hs, n, k.shape[0], _p(k, _ip), _p(acc, _ip), _p(info, _dp),
                                          _p(near, _ip))
    if rc != 0:
        msgs = {lib().pimc_chain_last_error(c.h).decode() for c in chains} - {""}
        raise PimcError(f"run_chains_device ({rc}): {'; '.join(sorted(msgs))}")
    return acc.astype(bool), info, near


def gen_worldlines(seed, side, M, N, lam, tau, kind=0):
    """Seeded worldlines [M][N][3] from the library's MTRand stream (host code)."""
    b = np.zeros((M, N, 3))
    rc_ = lib().pimc_gen_worldlines(seed, side, M, N, lam, tau, kind, _p(b, _dp))
    if rc_ != 0:
        raise PimcError(f"gen_worldlines failed ({rc_})")
    return b


def copy_probe(nbytes, iters=20):
    ms = np.zeros(1)
    rc_ = lib().pimc_copy_probe(nbytes, iters, _p(ms, _dp))
    if rc_ != 0:
        raise PimcError(f"copy probe failed ({rc_})")
    return float(ms[0])


class _PinnedBlock:
    """Page-locked host memory from pimc_host_alloc, freed with the last array
    viewing it."""

    def __init__(self, nbytes):
        p = C.c_void_p()
        rc = lib().pimc_host_alloc(nbytes, C.byref(p))
        if rc != 0:
            raise PimcError(f"pimc_host_alloc: {lib().pimc_last_error(None).decode()} ({rc})")
        self.ptr, self.nbytes = p.value, nbytes

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().pimc_host_free(self.ptr)
            self.ptr = None


def pinned_empty(shape, dtype=np.float64):
    """An uninitialised array in page-locked host memory (for large results such
    as F(q, tau)); the memory is released with the last array viewing it."""
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) if np.ndim(shape) else int(shape)
    blk = _PinnedBlock(max(1, n) * dt.itemsize)
    buf = (C.c_char * blk.nbytes).from_address(blk.ptr)
    buf._pimc_block = blk          # the ctypes buffer keeps the block alive
    return np.frombuffer(buf, dtype=dt, count=n).reshape(shape)


def device_synchronize():
    rc_ = lib().pimc_device_synchronize()
    if rc_ != 0:
        raise PimcError(lib().pimc_last_error(None).decode())
