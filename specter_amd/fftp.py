"""Host-side mirror of the reference's FFTP/FC-Gram, operator and boundary
interface, over the libspx C ABI.

Method names, argument meaning and in/out conventions follow the reference
(fftp/fftp.fpp, fftp/fcgram_mod.f90, pseudo/pseudospec_hd.f90,
boundary/boundary_mod.fpp, boundary/vboundary.f90).  Device arrays are torch
tensors on a ROCm device using C-order shapes that reproduce the Fortran
memory layout:

* spectral  Fortran (nz, ny, nkx)  ->  torch (nkx, ny, nz) complex128
* real      Fortran (nx, ny, nz)   ->  torch (nz, ny, nx)  float64

In mixed-precision plans (``precision=SPX_FP32_MIXED``) the outputs of
``gradre`` and ``advect`` are complex64; every other array is float64.

A batched plan (``members > 1``) holds independent runs of one grid: arrays
get a leading member axis, calls process the ``active`` leading members, and
the diagnostics return one value per active member.

Every call goes through libspx; there is no host fallback.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_double, c_void_p

import numpy as np

from . import lib as _l
from .tables import load_dirichlet_tables, load_neumann_tables


def _torch():
    import torch
    return torch


class FCPlan:
    """FCPLAN + fftp plans + the kes/grid module state (fcgram_mod.f90:34-144,
    specter.fpp:683-800) for one grid on one HIP stream."""

    def __init__(self, nx: int, ny: int, nz: int, Cz: int, oz: int, Lx: float = 1.0,
                 Ly: float = 1.0, Lz: float = 1.0, tdir: str | None = None, stream=None,
                 precision: int = _l.SPX_FP64, members: int = 1, Cy: int = 0, oy: int = 0):
        """Cy > 0: non-periodic y (FC-Gram continuation with Cy points and
        matching order oy, ny including the Cy points, Ly the physical length;
        fcgram_mod.f90:49-144 with C(2) = Cy)."""
        lib = _l.load_library()
        self._lib = lib
        g = _l.SpxGrid(nx, ny, nz, Cz, oz, Lx, Ly, Lz, precision)
        A = Q = Ay = Qy = None
        if Cz > 0:
            A, Q = load_dirichlet_tables(Cz, oz, tdir)
        if Cy > 0:
            Ay, Qy = load_dirichlet_tables(Cy, oy, tdir)
        self.A, self.Q, self.Ay, self.Qy = A, Q, Ay, Qy
        torch = _torch()
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        self.stream = stream
        ptr = c_void_p()

        def dp(a):
            return a.ctypes.data_as(ctypes.POINTER(c_double)) if a is not None else None

        _l.check(lib.spx_plan_create_fcy(byref(g), dp(A), dp(Q), int(Cy), int(oy), dp(Ay),
                                         dp(Qy), int(members), c_void_p(stream), byref(ptr)))
        self.members = int(members)
        self.active = int(members)
        self._p = ptr
        info = _l.SpxPlanInfo()
        _l.check(lib.spx_plan_get_info(ptr, byref(info)))
        self.nx, self.ny, self.nz, self.Cz, self.oz = nx, ny, nz, Cz, oz
        self.Cy, self.oy = Cy, oy
        # the Neumann reconstructors of the z walls (magnetic boundary
        # conditions), when the table directory holds them for this order
        self.neumann = False
        if Cz > 0:
            try:
                tabs = [load_neumann_tables(oz, o, tdir) for o in (1, 2)]
            except FileNotFoundError:
                tabs = None
            if tabs is not None:
                for o, (Qn, dxp) in zip((1, 2), tabs):
                    raw = np.concatenate([[dxp], np.asarray(Qn).ravel(order="F")])
                    _l.check(lib.spx_plan_load_neumann(ptr, o, dp(np.asfortranarray(Q)),
                                                       dp(np.ascontiguousarray(raw))))
                self.neumann = True
        self.Lx, self.Ly, self.Lz = Lx, Ly, Lz
        self.nkx, self.np = info.nkx, info.np
        self.z_fft_len, self.z_algorithm = info.z_fft_len, info.z_algorithm
        self.precision = precision

    # -- lifetime ---------------------------------------------------------------
    def close(self):
        if getattr(self, "_p", None):
            self._lib.spx_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._p

    def info(self) -> _l.SpxPlanInfo:
        info = _l.SpxPlanInfo()
        _l.check(self._lib.spx_plan_get_info(self._p, byref(info)))
        return info

    def sync(self):
        _l.check(self._lib.spx_plan_sync(self._p))

    # -- array helpers ------------------------------------------------------------
    def set_active(self, n: int):
        """Process the n leading members in every later call."""
        _l.check(self._lib.spx_plan_set_active(self._p, int(n)))
        self.active = int(n)

    @property
    def member_spec_shape(self):
        return (self.nkx, self.ny, self.nz)

    @property
    def member_real_shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def spec_shape(self):
        s = self.member_spec_shape
        return s if self.members == 1 else (self.members, *s)

    @property
    def real_shape(self):
        s = self.member_real_shape
        return s if self.members == 1 else (self.members, *s)

    def _out(self, vals):
        """scalar for single plans, one value per active member otherwise"""
        return vals[0] if self.members == 1 else np.array(vals[: self.active])

    def empty_spectral(self):
        torch = _torch()
        return torch.empty(self.spec_shape, dtype=torch.complex128, device="cuda")

    def empty_real(self):
        torch = _torch()
        return torch.empty(self.real_shape, dtype=torch.float64, device="cuda")

    @property
    def mixed(self) -> bool:
        return self.precision == _l.SPX_FP32_MIXED

    @property
    def nonlinear_dtype(self):
        """dtype of the gradre / advect results."""
        torch = _torch()
        return torch.complex64 if self.mixed else torch.complex128

    def empty_nonlinear(self):
        """gradre / advect result array (all members)."""
        torch = _torch()
        return torch.empty(self.spec_shape, dtype=self.nonlinear_dtype, device="cuda")

    def _spec(self, t, name, dtype=None):
        torch = _torch()
        dtype = dtype or torch.complex128
        if not (t.is_cuda and t.dtype == dtype and t.is_contiguous()
                and tuple(t.shape) == self.spec_shape):
            raise ValueError(f"{name}: expected contiguous cuda {dtype} {self.spec_shape}")
        return c_void_p(t.data_ptr())

    def _real(self, t, name):
        torch = _torch()
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
                and tuple(t.shape) == self.real_shape):
            raise ValueError(f"{name}: expected contiguous cuda float64 {self.real_shape}")
        return c_void_p(t.data_ptr())

    # -- transforms (fftp.fpp) ----------------------------------------------------------
    def fftp3d_real_to_complex(self, inp, out):
        """fftp.fpp:388-425."""
        _l.check(self._lib.spx_fftp3d_real_to_complex(self._p, self._real(inp, "in"),
                                                      self._spec(out, "out")))
        return out

    def fftp3d_complex_to_real(self, inp, out):
        """fftp.fpp:789-821 (the input may be overwritten, as in the reference)."""
        _l.check(self._lib.spx_fftp3d_complex_to_real(self._p, self._spec(inp, "in"),
                                                      self._real(out, "out")))
        return out

    def fftp2d_real_to_complex_xy(self, inp, out):
        """fftp.fpp:428-524."""
        _l.check(self._lib.spx_fftp2d_real_to_complex_xy(self._p, self._real(inp, "in"),
                                                         self._spec(out, "out")))
        return out

    def fftp2d_complex_to_real_xy(self, inp, out):
        """fftp.fpp:824-919."""
        _l.check(self._lib.spx_fftp2d_complex_to_real_xy(self._p, self._spec(inp, "in"),
                                                         self._real(out, "out")))
        return out

    def fftp1d_real_to_complex_z(self, inout):
        """fftp.fpp:720-786 (continuation + z FFT, in place)."""
        _l.check(self._lib.spx_fftp1d_real_to_complex_z(self._p, self._spec(inout, "inout")))
        return inout

    def fftp1d_complex_to_real_z(self, inout):
        """fftp.fpp:1060-1094 (z IFFT, in place)."""
        _l.check(self._lib.spx_fftp1d_complex_to_real_z(self._p, self._spec(inout, "inout")))
        return inout

    def fftp1d_real_to_complex_x(self, inp, out):
        """fftp.fpp:620-716 (x r2c of every plane, S layout)."""
        _l.check(self._lib.spx_fftp1d_real_to_complex_x(self._p, self._real(inp, "in"),
                                                        self._spec(out, "out")))
        return out

    def fftp2d_real_to_complex_yz(self, inout):
        """fftp.fpp:527-616 (y and z continuation + 2D yz FFT, in place)."""
        _l.check(self._lib.spx_fftp2d_real_to_complex_yz(self._p, self._spec(inout, "inout")))
        return inout

    def fftp2d_complex_to_real_yz(self, inout):
        """fftp.fpp:922-956 (2D yz IFFT, in place)."""
        _l.check(self._lib.spx_fftp2d_complex_to_real_yz(self._p, self._spec(inout, "inout")))
        return inout

    def fftp1d_complex_to_real_x(self, inp, out):
        """fftp.fpp:961-1056 (x c2r of every plane)."""
        _l.check(self._lib.spx_fftp1d_complex_to_real_x(self._p, self._spec(inp, "in"),
                                                        self._real(out, "out")))
        return out

    # -- MHD operators ---------------------------------------------------------------
    def curlk(self, a, b, c, dir: int):
        """c = component dir of curl from a, b (pseudospec_hd.f90:130-206)."""
        _l.check(self._lib.spx_curlk(self._p, self._spec(a, "a"), self._spec(b, "b"),
                                     self._spec(c, "c"), int(dir)))
        return c

    def prodre(self, a, b, c, d, e, f):
        """(d,e,f) = (curl A) x A (pseudospec_hd.f90:322-402)."""
        _l.check(self._lib.spx_prodre(self._p, *(self._spec(x, n) for x, n in
                                                 zip((a, b, c, d, e, f), "abcdef"))))
        return d, e, f

    def vector(self, a, b, c, d, e, f, g, h, i):
        """(g,h,i) = (a,b,c) x (d,e,f) (pseudospec_mhd.f90: vector)."""
        _l.check(self._lib.spx_vector(self._p, *(self._spec(x, n) for x, n in
                                                 zip((a, b, c, d, e, f, g, h, i),
                                                     "abcdefghi"))))
        return g, h, i

    def a_imposebc_and_project(self, ax, ay, az, ph, bczsta: int, bczend: int):
        """bboundary.f90:100-189 (wall kinds 0 conducting, 1 insulating)."""
        _l.check(self._lib.spx_a_imposebc_and_project(
            self._p, self._spec(ax, "ax"), self._spec(ay, "ay"), self._spec(az, "az"),
            self._spec(ph, "ph"), int(bczsta), int(bczend)))

    def fc_continue(self, inout):
        """fftp.fpp:755-773."""
        _l.check(self._lib.spx_fc_continue(self._p, self._spec(inout, "inout")))
        return inout

    # -- operators (pseudospec_hd.f90 / pseudospec_phd.f90) ------------------------------
    def derivk(self, a, b, dir: int):
        _l.check(self._lib.spx_derivk(self._p, self._spec(a, "a"), self._spec(b, "b"), dir))
        return b

    def laplak(self, a, b):
        _l.check(self._lib.spx_laplak(self._p, self._spec(a, "a"), self._spec(b, "b")))
        return b

    def fc_filter(self, a):
        _l.check(self._lib.spx_fc_filter(self._p, self._spec(a, "a")))
        return a

    def gradre(self, a, b, c, d, e, f):
        """d, e, f: ``nonlinear_dtype`` (complex64 in mixed plans)."""
        nd = self.nonlinear_dtype
        _l.check(self._lib.spx_gradre(self._p, self._spec(a, "a"), self._spec(b, "b"),
                                      self._spec(c, "c"), self._spec(d, "d", nd),
                                      self._spec(e, "e", nd), self._spec(f, "f", nd)))
        return d, e, f

    def advect(self, a, b, c, d, e):
        """e: ``nonlinear_dtype`` (complex64 in mixed plans)."""
        _l.check(self._lib.spx_advect(self._p, *(self._spec(t, n) for t, n in
                                                 zip((a, b, c, d), "abcd")),
                                      self._spec(e, "e", self.nonlinear_dtype)))
        return e

    # -- boundary conditions (boundary/) ----------------------------------------------
    def sol_project(self, a, b, c, d, bctarget=1, bczsta=0, bczend=0):
        _l.check(self._lib.spx_sol_project(self._p, self._spec(a, "a"), self._spec(b, "b"),
                                           self._spec(c, "c"), self._spec(d, "d"), bctarget,
                                           bczsta, bczend))

    def v_imposebc_and_project(self, vx, vy, vz, pr, rki: int, ord: int, v_zsta=(0.0, 0.0),
                               v_zend=(0.0, 0.0)):
        s = (c_double * 2)(*v_zsta)
        e = (c_double * 2)(*v_zend)
        _l.check(self._lib.spx_v_imposebc_and_project(
            self._p, self._spec(vx, "vx"), self._spec(vy, "vy"), self._spec(vz, "vz"),
            self._spec(pr, "pr"), rki, ord, s, e))

    def s_imposebc(self, th):
        _l.check(self._lib.spx_s_imposebc(self._p, self._spec(th, "th")))

    # -- diagnostics -----------------------------------------------------------------------
    def _vals(self):
        return (c_double * self.members)()

    def energy(self, a, b, c, kin: int):
        out = self._vals()
        _l.check(self._lib.spx_energy(self._p, self._spec(a, "a"), self._spec(b, "b"),
                                      self._spec(c, "c"), kin, out))
        return self._out(list(out))

    def cross(self, a, b, c, d, e, f):
        out = self._vals()
        _l.check(self._lib.spx_cross(self._p, *(self._spec(t, n) for t, n in
                                                zip((a, b, c, d, e, f), "abcdef")), out))
        return self._out(list(out))

    def helicity(self, a, b, c):
        """<v . curl v> over physical z (pseudospec_hd.f90:638-775)."""
        out = self._vals()
        _l.check(self._lib.spx_helicity(self._p, self._spec(a, "a"), self._spec(b, "b"),
                                        self._spec(c, "c"), out))
        return self._out(list(out))

    def divergence(self, a, b, c):
        out = self._vals()
        _l.check(self._lib.spx_divergence(self._p, self._spec(a, "a"), self._spec(b, "b"),
                                          self._spec(c, "c"), out))
        return self._out(list(out))

    def bouncheck_z(self, a, b=None):
        bot, top = self._vals(), self._vals()
        _l.check(self._lib.spx_bouncheck_z(self._p, self._spec(a, "a"),
                                           self._spec(b, "b") if b is not None else None,
                                           bot, top))
        return self._out(list(bot)), self._out(list(top))

    def variance(self, a, kin: int):
        """pseudospec_phd.f90:116-176."""
        out = self._vals()
        _l.check(self._lib.spx_variance(self._p, self._spec(a, "a"), kin, out))
        return self._out(list(out))

    def product(self, a, b):
        """pseudospec_phd.f90:178-235."""
        out = self._vals()
        _l.check(self._lib.spx_product(self._p, self._spec(a, "a"), self._spec(b, "b"), out))
        return self._out(list(out))

    def _per_member(self, vals, k):
        arr = np.array(vals).reshape(self.members, k)[: self.active]
        return tuple(arr[0]) if self.members == 1 else arr

    def wall_shear(self, a, b, nu: float):
        """(tau_x, tau_y) at z = 0 then z = Lz: nu <d(a, b)/dz> over the wall plane."""
        out = (c_double * (4 * self.members))()
        _l.check(self._lib.spx_wall_shear(self._p, self._spec(a, "a"), self._spec(b, "b"),
                                          c_double(nu), out))
        return self._per_member(list(out), 4)

    def nusselt(self, th):
        """Nusselt number at z = 0 and z = Lz for theta_total = 1 - z/Lz + th."""
        out = (c_double * (2 * self.members))()
        _l.check(self._lib.spx_nusselt(self._p, self._spec(th, "th"), out))
        return self._per_member(list(out), 2)

    def divergence_max(self, a, b, c):
        """max |div v| over the physical points, physical units."""
        out = self._vals()
        _l.check(self._lib.spx_divergence_max(self._p, self._spec(a, "a"), self._spec(b, "b"),
                                              self._spec(c, "c"), out))
        return self._out(list(out))

    def energy_spectrum(self, a, b, c, nbins: int):
        """Horizontal shell spectrum of <|v|^2> over physical z (nbins shells)."""
        out = (c_double * (nbins * self.members))()
        _l.check(self._lib.spx_energy_spectrum(self._p, self._spec(a, "a"), self._spec(b, "b"),
                                               self._spec(c, "c"), nbins, out))
        arr = np.array(list(out)).reshape(self.members, nbins)[: self.active]
        return arr[0] if self.members == 1 else arr

    # -- grid values (host copies of what the plan uses) --------------------------------
    def grid_z(self) -> np.ndarray:
        if self.Cz > 0:
            dz = self.Lz / (self.nz - self.Cz - 1)
        else:
            dz = self.Lz * 2.0 * 3.14159265358979323846 / self.nz
        return dz * np.arange(self.nz, dtype=np.float64)
