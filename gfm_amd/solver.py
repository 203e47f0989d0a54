"""Host mirror of the reference's SPHASE/GFLOW/ADLBL3 interface over the
HIP C-ABI (include/gfm_hip.h).

The reference drives the combustion space through argument-less Fortran
subroutines acting on module state (SPHASE C/Sphase.f90:10, GFLOW
C/Gflow.f90:6, ADLBL3 C/Adlbl3.f90:10).  ``CombustionSpace`` keeps that
shape: it owns one case's device state and exposes ``sphase``, ``gflow``,
``adlbl3`` and ``calc_residual`` with the reference's argument meaning.
Errors from the library raise ``GfmError`` (the reference's stop_run).
There is no CPU fallback: without the built library every call raises.
"""
import ctypes as ct
import os

import numpy as np

from .abi import GfmCtrl, GfmMsCtrl, GfmRadCtrl, GfmWallCtrl, host_arrays, ms_arrays, rad_arrays, wall_arrays

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgfm_hip.so")
_lib = None


class GfmError(RuntimeError):
    pass


def lib():
    """Load libgfm_hip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GfmError(f"{LIB_PATH} is missing: build it with `make -C gfm_amd/csrc` "
                           "(or __graft_entry__.build()); there is no CPU fallback")
        L = ct.CDLL(LIB_PATH)
        P = ct.c_void_p
        D = ct.POINTER(ct.c_double)
        L.gfm_hip_last_error.restype = ct.c_char_p
        L.gfm_hip_ctrl_size.restype = ct.c_size_t
        L.gfm_hip_init.argtypes = [ct.POINTER(GfmCtrl), P, ct.c_int, ct.POINTER(P)]
        L.gfm_hip_upload.argtypes = [P, ct.POINTER(GfmCtrl), P]
        L.gfm_hip_download.argtypes = [P, ct.POINTER(GfmCtrl), P]
        L.gfm_hip_initsv_tail.argtypes = [P, ct.POINTER(GfmCtrl)]
        L.gfm_hip_gflow.argtypes = [P, ct.POINTER(GfmCtrl)]
        L.gfm_hip_sphase.argtypes = [P, ct.POINTER(GfmCtrl), ct.c_int, D, ct.POINTER(ct.c_int)]
        L.gfm_hip_adlbl3.argtypes = [P, ct.POINTER(GfmCtrl), ct.c_int, D, D, D, D] + [ct.c_int] * 6 + [D]
        L.gfm_hip_calc_residual.argtypes = [P, D, D, D, D] + [ct.c_int] * 6 + [D]
        L.gfm_hip_finalize.argtypes = [P]
        L.gfm_hip_set_pguess.argtypes = [P, P]
        L.gfm_hip_pcorr_pair.argtypes = [P, ct.POINTER(ct.c_float), ct.POINTER(ct.c_float)]
        L.gfm_hip_plane_flows.argtypes = [P, D, D, D, D, D]
        L.gfm_hip_autotune.argtypes = [P, ct.POINTER(GfmCtrl), ct.c_int, ct.c_char_p, ct.c_int]
        L.gfm_hip_ms_upload.argtypes = [P, ct.POINTER(GfmMsCtrl), P]
        L.gfm_hip_ms_download.argtypes = [P, ct.POINTER(GfmMsCtrl), P]
        L.gfm_hip_initsvm.argtypes = [P, ct.POINTER(GfmMsCtrl), ct.c_int]
        L.gfm_hip_gslv3m.argtypes = [P, ct.POINTER(GfmMsCtrl)]
        L.gfm_hip_msflow.argtypes = [P, ct.POINTER(GfmMsCtrl)]
        L.gfm_hip_rad_emis.argtypes = [P, ct.POINTER(GfmRadCtrl)]
        L.gfm_hip_rad_qe_only.argtypes = [P, ct.POINTER(GfmRadCtrl)]
        L.gfm_hip_rad_download.argtypes = [P, ct.POINTER(GfmRadCtrl), P]
        L.gfm_hip_rad_upload.argtypes = [P, P]
        L.gfm_hip_rad_ctrl_size.restype = ct.c_size_t
        L.gfm_hip_wall_setup.argtypes = [P, ct.POINTER(GfmWallCtrl), P]
        L.gfm_hip_wallq.argtypes = [P, ct.POINTER(GfmWallCtrl)]
        L.gfm_hip_wall_download.argtypes = [P, ct.POINTER(GfmWallCtrl), P]
        L.gfm_hip_wall_upload.argtypes = [P, ct.POINTER(GfmWallCtrl), P]
        L.gfm_hip_wall_ctrl_size.restype = ct.c_size_t
        L.gfm_hip_kernel_times.argtypes = [P, ct.c_int, ct.c_char_p, ct.c_int, D, D, ct.POINTER(ct.c_int), ct.c_int,
                                           ct.POINTER(ct.c_int)]
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        msg = lib().gfm_hip_last_error().decode(errors="replace")
        raise GfmError(f"{what} failed (code {rc}): {msg}")


def _dp(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


class CombustionSpace:
    """Device-resident state of one case (gfm_hip_init .. gfm_hip_finalize)."""

    def __init__(self, case, device=0, initsv=True):
        self.case = case
        self._h = host_arrays(case.arrays)
        self._ctx = ct.c_void_p()
        _check(lib().gfm_hip_init(ct.byref(case.ctrl), ct.cast(ct.byref(self._h), ct.c_void_p), int(device),
                                  ct.byref(self._ctx)), "gfm_hip_init")
        if initsv:
            self.initsv_tail()

    # -- the reference's subroutine interface --
    def initsv_tail(self):
        _check(lib().gfm_hip_initsv_tail(self._ctx, ct.byref(self.case.ctrl)), "gfm_hip_initsv_tail")

    def gflow(self):
        """One GFLOW (one SIMPLER iteration) at itr_gas+1."""
        self.case.ctrl.itr_gas += 1
        _check(lib().gfm_hip_gflow(self._ctx, ct.byref(self.case.ctrl)), "gfm_hip_gflow")
        return self.case.ctrl.res_mass

    def sphase(self, max_iters):
        """SPHASE: iterate until res_mass < bgcon, itr_gend/maxgi, or
        max_iters; returns the res_mass history of the iterations run."""
        hist = np.zeros(max(int(max_iters), 1))
        n = ct.c_int(0)
        rc = lib().gfm_hip_sphase(self._ctx, ct.byref(self.case.ctrl), int(max_iters), _dp(hist), ct.byref(n))
        _check(rc, "gfm_hip_sphase")
        return hist[: n.value]

    def adlbl3(self, nelg, ap, as_, bs, xt, box):
        """ADLBL3(XT,IB,IE,JB,JE,KB,KE) on host arrays (XT updated in place)."""
        resg = np.zeros(2)
        _check(lib().gfm_hip_adlbl3(self._ctx, ct.byref(self.case.ctrl), int(nelg), _dp(ap), _dp(as_), _dp(bs), _dp(xt),
                                    *map(int, box), _dp(resg)), "gfm_hip_adlbl3")
        return resg

    def calc_residual(self, ap, as_, bs, xt, box):
        r = np.zeros(1)
        _check(lib().gfm_hip_calc_residual(self._ctx, _dp(ap), _dp(as_), _dp(bs), _dp(xt), *map(int, box), _dp(r)),
               "gfm_hip_calc_residual")
        return float(r[0])

    def autotune(self, iters=1):
        """gfm_hip_autotune: 6*iters ordinary iterations that time every
        launch shape of the per-position box kernels; returns the table of
        choices (text)."""
        buf = ct.create_string_buffer(8192)
        _check(lib().gfm_hip_autotune(self._ctx, ct.byref(self.case.ctrl), int(iters), buf, len(buf)), "gfm_hip_autotune")
        return buf.value.decode()

    # -- minor species (INITSVM, GSLV3M, MSFLOW) on a gfm_amd.minor.MinorSpecies --
    def ms_upload(self, ms):
        a = ms_arrays(ms.arrays)
        _check(lib().gfm_hip_ms_upload(self._ctx, ct.byref(ms.ctrl), ct.cast(ct.byref(a), ct.c_void_p)), "gfm_hip_ms_upload")

    def ms_download(self, ms):
        a = ms_arrays(ms.arrays)
        _check(lib().gfm_hip_ms_download(self._ctx, ct.byref(ms.ctrl), ct.cast(ct.byref(a), ct.c_void_p)),
               "gfm_hip_ms_download")
        return ms

    def initsvm(self, ms, restart=False):
        """INITSVM on the device: a new start from the gas field, or
        (restart) from the state of the last ms_upload."""
        _check(lib().gfm_hip_initsvm(self._ctx, ct.byref(ms.ctrl), int(bool(restart))), "gfm_hip_initsvm")

    def gslv3m(self, ms):
        _check(lib().gfm_hip_gslv3m(self._ctx, ct.byref(ms.ctrl)), "gfm_hip_gslv3m")

    def msflow(self, ms):
        _check(lib().gfm_hip_msflow(self._ctx, ct.byref(ms.ctrl)), "gfm_hip_msflow")
        return ms.ctrl.iterations

    # ---- radiation transport (RAD_EMIS / rad_qe_only) ----
    def rad_emis(self, rad, download=True):
        """One full radiation pass on the device state (the minor species
        must be on the device: ms_upload or initsvm first).  rad:
        gfm_amd.radiation.Radiation, whose ctrl receives the sums and, with
        download, whose arrays receive QE, QA, QAL and AKL."""
        _check(lib().gfm_hip_rad_emis(self._ctx, ct.byref(rad.ctrl)), "gfm_hip_rad_emis")
        return self.rad_download(rad) if download else rad

    def rad_qe_only(self, rad, download=True):
        """rad_qe_only: QE of every open cell (rad.ctrl.qeg0 its sum)."""
        _check(lib().gfm_hip_rad_qe_only(self._ctx, ct.byref(rad.ctrl)), "gfm_hip_rad_qe_only")
        return self.rad_download(rad) if download else rad

    def rad_download(self, rad):
        a = rad_arrays(rad.arrays)
        _check(lib().gfm_hip_rad_download(self._ctx, ct.byref(rad.ctrl), ct.cast(ct.byref(a), ct.c_void_p)),
               "gfm_hip_rad_download")
        return rad

    def rad_upload(self, rad):
        """QE and QA of rad.arrays onto the device (the fields of a saved
        radiation pass) for the enthalpy source GFM_RADSRC_FIELD."""
        a = rad_arrays(rad.arrays)
        _check(lib().gfm_hip_rad_upload(self._ctx, ct.cast(ct.byref(a), ct.c_void_p)), "gfm_hip_rad_upload")

    def wall_setup(self, wall, iwall=None):
        """prepare_for_wall_radiation_calculation on the device's cell codes
        (wall: gfm_amd.wall.Wall with its inputs in ctrl; iwall: the wall
        group of each cell or None); allocates and downloads wall's patch
        arrays."""
        iw = None if iwall is None else np.ascontiguousarray(iwall.ravel(order="F"), dtype=np.int16)
        _check(lib().gfm_hip_wall_setup(self._ctx, ct.byref(wall.ctrl),
                                        None if iw is None else iw.ctypes.data_as(ct.c_void_p)), "gfm_hip_wall_setup")
        wall.allocate(wall.ctrl.b_cnt)
        return self.wall_download(wall)

    def wallq(self, wall, download=True):
        """One radiation pass's wall exchange after rad_emis (wall.ctrl
        carries i1st_rad_done, iterw, the mean wall temperature and q_melt
        between passes and receives the sums)."""
        _check(lib().gfm_hip_wallq(self._ctx, ct.byref(wall.ctrl)), "gfm_hip_wallq")
        return self.wall_download(wall) if download else wall

    def wall_download(self, wall):
        a = wall_arrays(wall.arrays)
        _check(lib().gfm_hip_wall_download(self._ctx, ct.byref(wall.ctrl), ct.cast(ct.byref(a), ct.c_void_p)),
               "gfm_hip_wall_download")
        return wall

    def wall_upload(self, wall):
        """The patch state of a radiation record (radiosity, q_conv, b_t, ...)
        onto the device; the enthalpy sink takes the uploaded q_conv."""
        a = wall_arrays(wall.arrays)
        _check(lib().gfm_hip_wall_upload(self._ctx, ct.byref(wall.ctrl), ct.cast(ct.byref(a), ct.c_void_p)),
               "gfm_hip_wall_upload")
        return wall

    def set_pguess(self, net):
        """Attach a gfm_amd.pnet.PNet as the learned p' initial guess (None
        detaches); used when ctrl.pcorr_mode = 1 and ctrl.pcorr_guess = 1."""
        self._pnet = net
        _check(lib().gfm_hip_set_pguess(self._ctx, net.handle if net is not None else None), "gfm_hip_set_pguess")

    @property
    def cells(self):
        c = self.case.ctrl
        return (c.lp - 6) // 2 + 1, (c.np - 6) // 2 + 1, (c.mp - 6) // 2 + 1

    def pcorr_pair(self):
        """(x [nz,ny,nx,2], p [nz,ny,nx]) of the last p' solve run with
        ctrl.pcorr_dump = 1: x[...,0] = BS/AP, x[...,1] = open-row mask,
        p = p' after the solve."""
        nz, ny, nx = self.cells
        x = np.zeros((nz, ny, nx, 2), dtype=np.float32)
        p = np.zeros((nz, ny, nx), dtype=np.float32)
        _check(lib().gfm_hip_pcorr_pair(self._ctx, x.ctypes.data_as(ct.POINTER(ct.c_float)),
                                        p.ctypes.data_as(ct.POINTER(ct.c_float))), "gfm_hip_pcorr_pair")
        return x, p

    def plane_flows(self):
        """FLX1D's flow rates of the last iteration run with
        ctrl.want_plane_flows = 1 (C/Flx1d.f90:16-73): dict of (M, 2) arrays
        flx, fly, flyh, flux_z, fluxh_z, row I-1 for face plane I, columns
        + and - direction (kg/s; enthalpy flows in W)."""
        mp, np_, lp = self.case.shape
        out = {k: np.zeros((m, 2), order="F") for k, m in
               (("flx", mp), ("fly", np_), ("flyh", np_), ("flux_z", lp), ("fluxh_z", lp))}
        dp = ct.POINTER(ct.c_double)
        _check(lib().gfm_hip_plane_flows(self._ctx, *(out[k].ctypes.data_as(dp)
                                                      for k in ("flx", "fly", "flyh", "flux_z", "fluxh_z"))),
               "gfm_hip_plane_flows")
        return out

    def kernel_times(self, enable=-1, cap=128):
        """Per-kernel HIP-event timings: enable=1 resets and starts, 0 stops,
        -1 only reads.  Returns {name: (ms, algorithmic_bytes, launches)}."""
        nl = 64
        names = ct.create_string_buffer(nl * cap)
        ms = np.zeros(cap)
        by = np.zeros(cap)
        la = (ct.c_int * cap)()
        n = ct.c_int(0)
        _check(lib().gfm_hip_kernel_times(self._ctx, int(enable), names, nl, _dp(ms), _dp(by), la, cap, ct.byref(n)),
               "gfm_hip_kernel_times")
        out = {}
        for r in range(n.value):
            nm = names.raw[r * nl:(r + 1) * nl].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[r]), float(by[r]), int(la[r]))
        return out

    def download(self):
        """Copy the device state back into case.arrays (sav3f/sprint point)."""
        _check(lib().gfm_hip_download(self._ctx, None, ct.cast(ct.byref(self._h), ct.c_void_p)), "gfm_hip_download")
        return self.case

    def upload(self):
        _check(lib().gfm_hip_upload(self._ctx, ct.byref(self.case.ctrl), ct.cast(ct.byref(self._h), ct.c_void_p)),
               "gfm_hip_upload")

    def close(self):
        if self._ctx:
            lib().gfm_hip_finalize(self._ctx)
            self._ctx = ct.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
