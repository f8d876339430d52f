"""Device-resident HD / Boussinesq / rotating Boussinesq / MHD solver over libspx (the RK
loop of specter.fpp:1002-1165 with the include/hd, bouss, rotbouss and mhd stage bodies).

All state lives in HBM inside libspx; host code only sets it up, steps it and
reads diagnostics.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_void_p
from dataclasses import dataclass

import numpy as np

from . import lib as _l
from .fftp import FCPlan


class _DevView:
    """__cuda_array_interface__ view of a libspx device array (zero copy)."""

    def __init__(self, ptr: int, shape, typestr: str):
        self.__cuda_array_interface__ = {
            "shape": tuple(shape), "typestr": typestr, "data": (ptr, False),
            "version": 3, "strides": None}


@dataclass
class SolverParams:
    solver: str = "HD"        # "HD", "BOUSS", "ROTBOUSS" or "MHD"
    ord: int = 2              # RK order ORD
    dt: float = 1e-3
    nu: float = 2e-3
    kappa: float = 0.0
    xmom: float = 0.0         # gama*xmom (specter.fpp:480-491)
    xtemp: float = 0.0        # gama*xtemp
    v_zsta: tuple = (0.0, 0.0)
    v_zend: tuple = (0.0, 0.0)
    omega: tuple = (0.0, 0.0, 0.0)   # ROTBOUSS: (omegax, omegay, omegaz)
    mu: float = 0.0                  # MHD: magnetic diffusivity
    b0: tuple = (0.0, 0.0, 0.0)      # MHD: uniform field (bx0, by0, bz0)
    bzsta: int = 0                   # MHD: magnetic wall at z = 0 (0 conducting, 1 insulating)
    bzend: int = 0                   # MHD: magnetic wall at z = Lz

    def to_c(self) -> _l.SpxParams:
        p = _l.SpxParams()
        kinds = {"HD": _l.SPX_SOLVER_HD, "BOUSS": _l.SPX_SOLVER_BOUSS,
                 "ROTBOUSS": _l.SPX_SOLVER_ROTBOUSS, "MHD": _l.SPX_SOLVER_MHD}
        if self.solver.upper() not in kinds:
            raise ValueError(f"solver must be one of {sorted(kinds)}, not {self.solver!r}")
        p.solver = kinds[self.solver.upper()]
        p.ord, p.dt, p.nu, p.kappa = self.ord, self.dt, self.nu, self.kappa
        p.xmom, p.xtemp = self.xmom, self.xtemp
        p.v_zsta[0], p.v_zsta[1] = self.v_zsta
        p.v_zend[0], p.v_zend[1] = self.v_zend
        p.omega[0], p.omega[1], p.omega[2] = self.omega
        p.mu = self.mu
        p.b0[0], p.b0[1], p.b0[2] = self.b0
        p.bzsta, p.bzend = int(self.bzsta), int(self.bzend)
        return p


class Solver:
    """One run: an FCPlan plus a libspx solver state."""

    FIELDS = ("vx", "vy", "vz", "pr", "fx", "fy", "fz", "th", "fs",
              "ax", "ay", "az", "ph", "mx", "my", "mz")

    def __init__(self, plan: FCPlan, params: SolverParams):
        self.plan = plan
        self.params = params
        self._lib = plan._lib
        st = c_void_p()
        _l.check(self._lib.spx_state_create(plan.handle, byref(params.to_c()), byref(st)))
        self._s = st
        self.bouss = params.solver.upper() in ("BOUSS", "ROTBOUSS")

    @property
    def fields(self) -> tuple:
        """The state fields this solver kind holds (a subset of FIELDS)."""
        kind = self.params.solver.upper()
        base = ("vx", "vy", "vz", "pr", "fx", "fy", "fz")
        if kind in ("BOUSS", "ROTBOUSS"):
            return base + ("th", "fs")
        if kind == "MHD":
            return base + ("ax", "ay", "az", "ph", "mx", "my", "mz")
        return base

    def close(self):
        if getattr(self, "_s", None):
            self._lib.spx_state_destroy(self._s)
            self._s = None
        if getattr(self, "_single", None) is not None:
            self._single.close()
            self._single = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field(self, name: str):
        """Zero-copy torch view (complex128, (nkx, ny, nz), with a leading
        member axis for batched plans) of a state field."""
        import torch
        ptr = c_void_p()
        _l.check(self._lib.spx_state_field(self._s, name.encode(), byref(ptr)))
        return torch.as_tensor(_DevView(ptr.value, self.plan.spec_shape, "<c16"), device="cuda")

    def _member_plan(self) -> FCPlan:
        """A single plan of the same grid on the same stream, for the I/O of
        one member of a batched state."""
        if getattr(self, "_single", None) is None:
            p = self.plan
            self._single = FCPlan(p.nx, p.ny, p.nz, p.Cz, p.oz, p.Lx, p.Ly, p.Lz,
                                  stream=p.stream, precision=p.precision)
        return self._single

    # -- set-up -------------------------------------------------------------------
    def set_real_field(self, name: str, real, member: int | None = None):
        """field <- fftp3d_real_to_complex(real) (restart path, specter.fpp:894-899).
        real: numpy array or cuda tensor (nz, ny, nx).  For batched plans,
        member selects the one run to set."""
        import torch
        if isinstance(real, torch.Tensor):
            r = real.to(device="cuda", dtype=torch.float64).contiguous()
        else:
            r = torch.as_tensor(np.ascontiguousarray(real, dtype=np.float64)).cuda()
        member = self._single_member(member)
        if member is None:
            self.plan.fftp3d_real_to_complex(r, self.field(name))
        else:
            self._member_plan().fftp3d_real_to_complex(r, self.field(name)[member])
        self.plan.sync()

    def _single_member(self, member):
        """member of a one-member plan: the whole field (member must be 0)."""
        if member is not None and self.plan.members == 1:
            if member != 0:
                raise ValueError("member out of range for a one-member plan")
            return None
        return member

    def member_field(self, name: str, member: int):
        """Zero-copy view of one member's spectral field."""
        f = self.field(name)
        return f if self._single_member(member) is None else f[member]

    def set_member(self, member: int, nu: float, kappa: float = 0.0):
        """Diffusivities of one member of a batched state."""
        _l.check(self._lib.spx_state_set_member(self._s, member, nu, kappa))

    def copy_member(self, dst: int, src: int):
        _l.check(self._lib.spx_state_copy_member(self._s, dst, src))

    # -- stepping --------------------------------------------------------------------
    def rkstep1(self):
        _l.check(self._lib.spx_rkstep1(self.plan.handle, self._s))

    def substep(self, o: int):
        _l.check(self._lib.spx_substep(self.plan.handle, self._s, o))

    def step(self, nsteps: int = 1):
        """nsteps full RK steps (one captured HIP graph per step unless
        use_graph(False) was called)."""
        _l.check(self._lib.spx_step(self.plan.handle, self._s, nsteps))

    def use_graph(self, on: bool = True):
        _l.check(self._lib.spx_state_set_graph(self._s, int(on)))

    # -- diagnostics (hd_global.f90 / vdiagnostic) -----------------------------------------
    def hdcheck(self):
        p = self.plan
        vx, vy, vz = self.field("vx"), self.field("vy"), self.field("vz")
        fx, fy, fz = self.field("fx"), self.field("fy"), self.field("fz")
        return (p.energy(vx, vy, vz, 1), p.energy(vx, vy, vz, 0),
                p.cross(vx, vy, vz, fx, fy, fz))

    def mhdcheck(self):
        """The values mhdcheck writes with hel = crs = 1 (pseudospec_mhd.f90
        mhdcheck): (eng, ens, cur) of balance.txt, (engk, engm) of energy.txt,
        (helk, helm) of helicity.txt, (crh, asq) of cross.txt."""
        p = self.plan
        v = [self.field(n) for n in ("vx", "vy", "vz")]
        a = [self.field(n) for n in ("ax", "ay", "az")]
        engk, ens = p.energy(*v, 1), p.energy(*v, 0)
        engm, cur = p.energy(*a, 0), p.energy(*a, 2)
        helk, helm = p.helicity(*v), p.helicity(*a)
        asq = p.energy(*a, 1)
        d = [p.empty_spectral() for _ in range(3)]
        for k, (src, dst) in enumerate(zip(a, d), start=1):
            p.derivk(src, dst, k)
        crh = p.cross(*v, *d)
        return (engk + engm, ens, cur, engk, engm, helk, helm, crh, asq)

    def helicity(self):
        """The helicity.txt value of hdcheck with hel=1 (pseudospec_hd.f90:980-1003)."""
        return self.plan.helicity(self.field("vx"), self.field("vy"), self.field("vz"))

    def vdiagnostic(self):
        p = self.plan
        vx, vy, vz = self.field("vx"), self.field("vy"), self.field("vz")
        dv = p.divergence(vx, vy, vz)
        tb, tt = p.bouncheck_z(vx, vy)
        nb, nt = p.bouncheck_z(vz)
        return dv, tb, tt, nb, nt

    def pscheck(self):
        """scalar.txt quantities (pseudospec_phd.f90:237-262)."""
        p = self.plan
        th, fs = self.field("th"), self.field("fs")
        return p.variance(th, 1), p.variance(th, 0), p.product(th, fs)

    def wall_diagnostics(self):
        """Wall shear (nu <du_t/dz> at both walls), max |div v| and, for the
        Boussinesq solver, the Nusselt number at both walls."""
        p = self.plan
        vx, vy, vz = self.field("vx"), self.field("vy"), self.field("vz")
        out = {"wall_shear": p.wall_shear(vx, vy, self.params.nu),
               "divergence_max": p.divergence_max(vx, vy, vz)}
        if self.bouss:
            out["nusselt"] = p.nusselt(self.field("th"))
        return out

    def real_field(self, name: str, member: int | None = None) -> np.ndarray:
        """Physical-space field (c2r / (nx ny nz)), physical planes only."""
        p = self.plan
        src = self.field(name)
        member = self._single_member(member)
        if member is not None:
            src, p = src[member], self._member_plan()
        elif p.members > 1:
            raise ValueError("real_field of a batched state needs a member")
        tmp = src.clone()
        out = p.empty_real()
        p.fftp3d_complex_to_real(tmp, out)
        p.sync()
        return (out[: p.np] / (p.nx * p.ny * p.nz)).cpu().numpy()

    def spectral(self, name: str, member: int | None = None) -> np.ndarray:
        t = self.field(name)
        member = self._single_member(member)
        if member is not None:
            t = t[member]
        self.plan.sync()
        return t.cpu().numpy()
