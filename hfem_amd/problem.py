"""Seeded synthetic HDG Navier-Stokes cases (the benchmark and parity inputs).

A ``Problem`` bundles what the reference's element path reads from
``app.bin``/``mesh.bin``/``sol.bin``: physics parameters
(``physicsparam = [gamma, eps_lm, Re, Pr, Minf, tau]``, preprocessing/digasopre.m:513),
the freestream boundary state ``ui`` (``app.bcs``, boundary type 1 = far field,
application/FM/fbou.cpp:126-217), the mesh arrays and the state ``UDG``, ``UH``,
``SH`` in the reference's flat layouts:

* ``UDG[ie*npv*nc + c*npv + n]``, components ``[u_0..u_ncu-1, q_x.., q_y.., (q_z..)]``
  (assembly/elementassembly.cpp:264,268),
* ``UH[g*nch + k]`` (assembly/elementassembly.cpp:605-609),
* ``dgnodes[ie*npv*ncd + x*npv + n]`` (assembly/elementassembly.cpp:39).

The state is freestream (rho=1, velocity along x, p = 1/(gamma Minf^2)) plus a
seeded low-mode sine perturbation of amplitude ``amp`` on every u component and
i.i.d. uniform[-qamp, qamp] values of q; ``UH`` is the face average of the
adjacent element traces (utilities/inituhat.m:1-27).
"""
from dataclasses import dataclass

import numpy as np

from . import mesh as meshmod
from .master import make_master

__all__ = ["Problem", "make_problem", "CONFIGS", "freestream_state"]


def freestream_state(nd, gamma, minf, angle=0.0, elevation=0.0):
    """rho = 1, |v| = 1 at `angle` (radians) from x in the x-y plane and `elevation`
    out of it (3D), p = 1/(gamma Minf^2)."""
    pinf = 1.0 / (gamma * minf ** 2)
    u = np.zeros(nd + 2)
    u[0] = 1.0
    ce = np.cos(elevation) if nd == 3 else 1.0
    u[1] = np.cos(angle) * ce
    u[2] += np.sin(angle) * ce
    if nd == 3:
        u[3] += np.sin(elevation)
    u[nd + 1] = 0.5 + pinf / (gamma - 1.0)
    return u


@dataclass
class Problem:
    name: str
    mesh: object
    nd: int
    ncu: int
    nch: int
    nc: int
    param: np.ndarray        # physicsparam [gamma, eps_lm, Re, Pr, Minf, tau]
    ui: np.ndarray           # freestream boundary state (nch,)
    UDG: np.ndarray          # (ne, nc, npv)
    UH: np.ndarray           # (nf*npf, nch)
    SH: np.ndarray           # (ne, nc, npv)
    fc_u: float = 0.0
    fc_q: float = 1.0
    tdep: int = 0
    time: float = 0.0

    @property
    def master(self):
        return self.mesh.master

    @property
    def ne(self):
        return self.mesh.ne

    @property
    def dims(self):
        m = self.master
        na = m.npv * self.ncu
        nb = m.ndf * self.nch
        return dict(nd=self.nd, ncu=self.ncu, nch=self.nch, nc=self.nc, npv=m.npv, ngv=m.ngv, npf=m.npf,
                    ngf=m.ngf, nfe=m.nfe, ndf=m.ndf, na=na, nb=nb, ne=self.ne, nf=self.mesh.nf)

    def dgnodes_ref(self):
        """(ne, ncd, npv) contiguous: the reference's dgnodes layout."""
        return np.ascontiguousarray(self.mesh.dgnodes.transpose(0, 2, 1))


def _uhat_average(mesh, UDG, ncu):
    m = mesh.master
    ne, nfe, npf = mesh.elcon.shape
    nsiz = mesh.nf * npf
    UH = np.zeros((nsiz, ncu))
    cnt = np.zeros(nsiz)
    perm = m.perm                                    # (npf, nfe)
    for a in range(nfe):
        g = mesh.elcon[:, a, :]                      # (ne, npf)
        vals = UDG[:, :ncu, :][:, :, perm[:, a]]     # (ne, ncu, npf)
        np.add.at(UH, g.reshape(-1), vals.transpose(0, 2, 1).reshape(-1, ncu))
        np.add.at(cnt, g.reshape(-1), 1.0)
    return UH / cnt[:, None]


def taylor_green_state(x, gamma, minf, amp):
    """Conservative state of the Taylor-Green vortex on the unit cube (period 1):
    rho = 1, velocity amp (sin X cos Y cos Z, -cos X sin Y cos Z, 0) with
    X = 2 pi x, ..., pressure 1/(gamma minf^2) + amp^2/16 (cos 2X + cos 2Y)(cos 2Z + 2).
    x: (..., 3) points; returns (..., 5)."""
    X, Y, Z = (2 * np.pi * x[..., i] for i in range(3))
    u = amp * np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -amp * np.cos(X) * np.sin(Y) * np.cos(Z)
    w = np.zeros_like(u)
    p = 1.0 / (gamma * minf ** 2) + amp ** 2 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    rho = np.ones_like(u)
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, E], axis=-1)


def make_problem(kind, ncell, porder, seed=0, tau=2.0, re=1000.0, pr=0.72, minf=0.3, gamma=1.4,
                 amp=1e-2, qamp=1e-2, jitter=0.05, angle=0.0, elevation=0.0, name=None, state="freestream",
                 tg_amp=0.1):
    """Build a seeded case.  kind: 'tri' (2D), 'tet' (3D), 'hex' (3D periodic).

    ``angle`` turns the freestream in the x-y plane; at angle 0 the flow is
    tangential to the y-faces of the box, where the normal velocity crosses zero
    and |A_n| of the far-field flux is not differentiable.  ``state`` is
    'freestream' (the freestream plus a seeded low-mode perturbation of amplitude
    ``amp``) or 'taylor_green' (3D only: the Taylor-Green vortex of velocity
    amplitude ``tg_amp`` plus seeded uniform noise of amplitude ``amp``)."""
    if kind == "tri":
        msh = meshmod.square_tri_mesh(ncell, porder, seed=seed, jitter=jitter)
    elif kind == "tet":
        msh = meshmod.cube_tet_mesh(ncell, porder, seed=seed, jitter=jitter)
    elif kind == "hex":
        msh = meshmod.cube_hex_mesh(ncell, porder, seed=seed, jitter=jitter, periodic=True)
    else:
        raise ValueError(kind)
    nd = msh.nd
    ncu = nd + 2
    nc = ncu * (nd + 1)
    m = make_master(nd, msh.elemtype, porder)
    rng = np.random.default_rng(seed + 7919)
    u0 = freestream_state(nd, gamma, minf, angle, elevation)
    x = msh.dgnodes                                   # (ne, npv, nd)
    ne = msh.ne
    UDG = np.zeros((ne, nc, m.npv))
    if state == "taylor_green":
        if nd != 3:
            raise ValueError("the Taylor-Green state is 3D")
        tg = taylor_green_state(x, gamma, minf, tg_amp)          # (ne, npv, 5)
        UDG[:, :ncu, :] = np.transpose(tg, (0, 2, 1)) + rng.uniform(-amp, amp, size=(ne, ncu, m.npv))
    elif state == "freestream":
        for c in range(ncu):
            pert = np.zeros(x.shape[:2])
            for _ in range(3):
                kvec = rng.integers(1, 3, size=nd)
                phase = rng.uniform(0, 2 * np.pi)
                pert += np.sin(2 * np.pi * (x @ kvec) + phase)
            UDG[:, c, :] = u0[c] + amp * pert / 3.0
    else:
        raise ValueError(state)
    UDG[:, ncu:, :] = rng.uniform(-qamp, qamp, size=(ne, nc - ncu, m.npv))
    UH = _uhat_average(msh, UDG, ncu)
    param = np.array([gamma, 0.0, re, pr, minf, tau])
    return Problem(name=name or f"{kind}{ncell}p{porder}", mesh=msh, nd=nd, ncu=ncu, nch=ncu, nc=nc,
                   param=param, ui=u0.copy(), UDG=UDG, UH=UH, SH=np.zeros_like(UDG))


# The configurations of the benchmark (name -> make_problem arguments).
CONFIGS = {
    "c1_tri_p2_16": dict(kind="tri", ncell=16, porder=2),
    "c2_tri_p4_128": dict(kind="tri", ncell=128, porder=4),
    "c3_tet_p3_14": dict(kind="tet", ncell=14, porder=3),
    "c5_hex_p2_24": dict(kind="hex", ncell=24, porder=2),
}
