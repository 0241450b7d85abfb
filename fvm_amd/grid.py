"""O-grid generator, seeded initial condition and the reference's grid / restart file formats.

The reference solver ships no grid generator and no random initial condition; it reads

* a node file ``NNNxNNNxNNN.dat`` (3D, one ``x y z`` record per node, K outer, J, I inner;
  ``LES/v2.3_LES-Reynolds stresses/grid.f90:46-67``) or ``NNNxNNN.dat`` (2D, ``x y`` records,
  ``cylinder/cylinder_v3.0.3_AF_original 4step_face_nl/grid.f90:28-43``), and
* a restart file ``RESTART000.DAT`` (3D, ``E24.16`` records; written at
  ``LES/.../writedata.f90:193-216``, read at ``LES/.../init.f90:183-208``) or ``RESTART.DAT``
  (2D, list-directed; written at ``cylinder/.../writedata.f90:77-90``, read at
  ``cylinder/.../init.f90:46-68``).

This module writes and reads exactly those layouts so the same files feed the reference build
(the oracle) and the GPU solver.  Arrays are numpy arrays in Fortran order, indexed
``[i, j, k]`` with ghosts: cell fields are ``(NI+1, NJ+1, NK+1)`` (indices 0..NI etc.), node
fields are ``(NI, NJ, NK)`` holding nodes 1..NI.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np

SEED = 20261202


def stretch_ratio(n_cells: int, h1: float, r_in: float, r_out: float) -> float:
    """Geometric ratio q with r_in + h1 (q^n - 1)/(q - 1) = r_out (bisection on q > 1)."""
    target = (r_out - r_in) / h1
    lo, hi = 1.0 + 1e-12, 2.0
    while (hi ** n_cells - 1.0) / (hi - 1.0) < target:
        hi *= 1.5
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if (mid ** n_cells - 1.0) / (mid - 1.0) < target:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


@dataclass
class OGrid:
    """Body-fitted O-grid around a unit-diameter cylinder.

    Node counts follow the reference's NI (radial), NJ (circumferential, node NJ coincides
    with node 1 at theta = 2 pi), NK (span).  ``nk == 1`` gives a 2D grid.
    """
    ni: int
    nj: int
    nk: int
    h1: float
    r_out: float
    lz: float = math.pi

    @property
    def q(self) -> float:
        return stretch_ratio(self.ni - 1, self.h1, 0.5, self.r_out)

    def nodes(self):
        """Node coordinates xg, yg, zg of shape (NI, NJ, NK) (Fortran order)."""
        q = self.q
        i = np.arange(self.ni, dtype=np.float64)
        r = 0.5 + self.h1 * (q ** i - 1.0) / (q - 1.0)
        r[-1] = self.r_out
        th = 2.0 * math.pi * np.arange(self.nj, dtype=np.float64) / (self.nj - 1)
        rr, tt = np.meshgrid(r, th, indexing="ij")
        x2 = rr * np.cos(tt)
        y2 = rr * np.sin(tt)
        # Node NJ is node 1 (theta = 2 pi): copy so the branch cut closes exactly.
        x2[:, -1] = x2[:, 0]
        y2[:, -1] = y2[:, 0]
        nk = max(self.nk, 1)
        xg = np.asfortranarray(np.repeat(x2[:, :, None], nk, axis=2))
        yg = np.asfortranarray(np.repeat(y2[:, :, None], nk, axis=2))
        if self.nk > 1:
            z = self.lz * np.arange(self.nk, dtype=np.float64) / (self.nk - 1)
        else:
            z = np.zeros(1)
        zg = np.asfortranarray(np.broadcast_to(z[None, None, :], xg.shape).copy())
        return xg, yg, zg


def grid_filename(ni: int, nj: int, nk: int | None = None) -> str:
    """File name the reference derives from the node counts (I3.3 fields)."""
    if nk is None:
        return f"{ni:03d}x{nj:03d}.dat"
    return f"{ni:03d}x{nj:03d}x{nk:03d}.dat"


def write_grid3d(path: str, xg, yg, zg) -> None:
    """One 'x y z' record per node, K outer, J middle, I inner (list-directed read)."""
    flat = np.stack([xg.ravel(order="F"), yg.ravel(order="F"), zg.ravel(order="F")], axis=1)
    np.savetxt(path, flat, fmt="%.17e")


def write_grid2d(path: str, xg, yg) -> None:
    flat = np.stack([xg[:, :, 0].ravel(order="F"), yg[:, :, 0].ravel(order="F")], axis=1)
    np.savetxt(path, flat, fmt="%.17e")


# --------------------------------------------------------------------------------------
# Initial condition
# --------------------------------------------------------------------------------------

def normal_stream(component: int, n: int, sigma: float, seed: int = SEED) -> np.ndarray:
    """N(0, sigma) draws from the Philox-4x32 counter-based generator; stream = component id."""
    gen = np.random.Generator(np.random.Philox(np.random.SeedSequence([seed, component])))
    return gen.normal(0.0, sigma, size=n)


def seeded_ic(ni: int, nj: int, nk: int, sigma: float, seed: int = SEED, three_d: bool = True):
    """Uniform flow u=1 plus N(0, sigma) on every interior cell of every velocity component.

    Returns dict of cell arrays u, v, w, p, t, nlx, nly, nlz, adv of shape (NI+1, NJ+1, NK+1)
    (3D) or (NI+1, NJ+1, 1) holding the single 2D plane.  Ghost cells: no-slip mirror at the
    wall (u(0) = -u(1)), free-stream mirror at the far field (u(NI) = 2 u_inf - u(NI-1)),
    periodic copies across the branch cut (J) and the span (K).
    """
    if three_d:
        shape = (ni + 1, nj + 1, nk + 1)
        interior = (slice(1, ni), slice(1, nj), slice(1, nk))
        n_int = (ni - 1) * (nj - 1) * (nk - 1)
        comps = ("u", "v", "w")
    else:
        shape = (ni + 1, nj + 1, 1)
        interior = (slice(1, ni), slice(1, nj), slice(0, 1))
        n_int = (ni - 1) * (nj - 1)
        comps = ("u", "v")
    out = {}
    for c, name in enumerate(comps):
        a = np.zeros(shape, order="F")
        base = 1.0 if name == "u" else 0.0
        noise = normal_stream(c, n_int, sigma, seed).reshape(
            a[interior].shape, order="F")
        a[interior] = base + noise
        # wall and far-field ghosts
        a[0] = -a[1]
        a[ni] = 2.0 * base - a[ni - 1]
        # periodic ghosts (branch cut, then span), interior rows of I included
        a[:, 0] = a[:, nj - 1]
        a[:, nj] = a[:, 1]
        if three_d:
            a[:, :, 0] = a[:, :, nk - 1]
            a[:, :, nk] = a[:, :, 1]
        out[name] = a
    if not three_d:
        out["w"] = np.zeros(shape, order="F")
    for name in ("p", "t", "nlx", "nly", "nlz", "adv"):
        out[name] = np.zeros(shape, order="F")
    return out


def seeded_temperature(ni: int, nj: int, nk: int, sigma: float, seed: int = SEED) -> np.ndarray:
    """Temperature start field: N(0, sigma) on the interior cells (stream 3 of the seed), wall
    ghost mirrored about T_WALL = 1 (T(0) = 2 - T(1)), far-field ghost about TIN = 0, periodic
    copies across the branch cut and the span."""
    t = np.zeros((ni + 1, nj + 1, nk + 1), order="F")
    interior = (slice(1, ni), slice(1, nj), slice(1, nk))
    t[interior] = normal_stream(3, (ni - 1) * (nj - 1) * (nk - 1), sigma, seed).reshape(
        t[interior].shape, order="F")
    t[0] = 2.0 - t[1]
    t[ni] = -t[ni - 1]
    t[:, 0] = t[:, nj - 1]
    t[:, nj] = t[:, 1]
    t[:, :, 0] = t[:, :, nk - 1]
    t[:, :, nk] = t[:, :, 1]
    return t


# --------------------------------------------------------------------------------------
# Restart files
# --------------------------------------------------------------------------------------

def fortran_e24_16(v: float) -> str:
    """Format like Fortran ``E24.16``: 0.dddddddddddddddd E+xx, right-justified in 24 chars."""
    v = float(v)
    if v == 0.0:
        body = ("-" if math.copysign(1.0, v) < 0 else "") + "0.0000000000000000E+00"
        return body.rjust(24)
    s = f"{abs(v):.15e}"  # d.ddddddddddddddde+XX (16 significant digits)
    mant, exp = s.split("e")
    digits = mant.replace(".", "")
    e = int(exp) + 1
    if abs(e) < 100:
        es = f"E{'+' if e >= 0 else '-'}{abs(e):02d}"
    else:
        es = f"{'+' if e >= 0 else '-'}{abs(e):03d}"
    body = ("-" if v < 0 else "") + "0." + digits + es
    return body.rjust(24)


def write_restart3d(path: str, f: dict, re: float, pe: float, dtime: float, rtime: float,
                    niter: int = 0) -> None:
    """RESTARTnnn.DAT layout: header (RE, PE, DTIME, RTIME, NITER_TIME), then per cell, I fastest,
    one record NLX NLY NLZ ADV and one record U V W P T, all E24.16 (TR1 separators)."""
    fmt = np.vectorize(fortran_e24_16, otypes=[object])
    a4 = np.stack([f[k].ravel(order="F") for k in ("nlx", "nly", "nlz", "adv")], axis=1)
    a5 = np.stack([f[k].ravel(order="F") for k in ("u", "v", "w", "p", "t")], axis=1)
    s4 = fmt(a4)
    s5 = fmt(a5)
    with open(path, "w") as fh:
        fh.write(" " + " ".join(fortran_e24_16(x) for x in (re, pe, dtime, rtime))
                 + " " + f"{niter:16d}" + "\n")
        lines = []
        for r4, r5 in zip(s4, s5):
            lines.append(" " + " ".join(r4) + "\n")
            lines.append(" " + " ".join(r5) + "\n")
        fh.writelines(lines)


def read_restart3d(path: str, ni: int, nj: int, nk: int) -> dict:
    """Inverse of write_restart3d (any whitespace-separated real format is accepted)."""
    with open(path) as fh:
        head = fh.readline().split()
        body = np.array(fh.read().split(), dtype=np.float64)
    n = (ni + 1) * (nj + 1) * (nk + 1)
    body = body.reshape(n, 9)
    out = {"re": float(head[0]), "pe": float(head[1]), "dtime": float(head[2]),
           "rtime": float(head[3]), "niter": int(head[4])}
    for c, name in enumerate(("nlx", "nly", "nlz", "adv", "u", "v", "w", "p", "t")):
        out[name] = body[:, c].reshape((ni + 1, nj + 1, nk + 1), order="F").copy(order="F")
    return out


def write_restart2d(path: str, f: dict, re: float, pr: float, dtime: float, vel_rot: float,
                    rtime: float = 0.0, niter: int = 0) -> None:
    """RESTART.DAT (2D, list-directed): RE PR DTIME VEL_ROT / RTIME NITER / per cell
    NLX NLY ADV and U V P T, I fastest over the single plane incl. ghosts."""
    with open(path, "w") as fh:
        fh.write(f" {re!r} {pr!r} {dtime!r} {vel_rot!r}\n")
        fh.write(f" {rtime!r} {niter}\n")
        a3 = np.stack([f[k][:, :, 0].ravel(order="F") for k in ("nlx", "nly", "adv")], axis=1)
        a4 = np.stack([f[k][:, :, 0].ravel(order="F") for k in ("u", "v", "p", "t")], axis=1)
        lines = []
        for r3, r4 in zip(a3, a4):
            lines.append(" " + " ".join(repr(float(x)) for x in r3) + "\n")
            lines.append(" " + " ".join(repr(float(x)) for x in r4) + "\n")
        fh.writelines(lines)


def read_restart2d(path: str, ni: int, nj: int) -> dict:
    with open(path) as fh:
        h1 = fh.readline().split()
        h2 = fh.readline().split()
        body = np.array(fh.read().split(), dtype=np.float64).reshape(-1, 7)
    out = {"re": float(h1[0]), "pr": float(h1[1]), "dtime": float(h1[2]),
           "vel_rot": float(h1[3]), "rtime": float(h2[0]), "niter": int(h2[1])}
    for c, name in enumerate(("nlx", "nly", "adv", "u", "v", "p", "t")):
        out[name] = body[:, c].reshape((ni + 1, nj + 1, 1), order="F").copy(order="F")
    return out


# --------------------------------------------------------------------------------------
# Named workloads (node counts as the reference's NI x NJ x NK)
# --------------------------------------------------------------------------------------

WORKLOADS = {
    # 2D laminar cylinder, Re=100: 129 radial x 257 circumferential nodes
    "cyl2d_re100": dict(ni=129, nj=257, nk=1, h1=0.002, r_out=30.0, re=100.0, dt=0.005,
                         sigma=1e-3),
    # 3D dynamic-Smagorinsky LES, Re=3900: 192 x 256 x 64 cells
    "les3d_re3900": dict(ni=193, nj=257, nk=65, h1=0.001, r_out=25.0, re=3900.0, dt=0.002,
                         sigma=0.05),
    # the same LES on 384 x 512 x 128 cells
    "les3d_re3900_25m": dict(ni=385, nj=513, nk=129, h1=0.001, r_out=25.0, re=3900.0,
                             dt=0.002, sigma=0.05),
    # the LES on a coarse 96 x 128 x 16 cell grid (wall spacing 0.002 D): long shedding runs
    # (stable over tens of thousands of steps at dt = 0.001)
    "les3d_re3900_coarse": dict(ni=97, nj=129, nk=17, h1=0.002, r_out=25.0, re=3900.0,
                                dt=0.001, sigma=0.05),
    # differentiable 2D case, Re=150: 97 x 193 nodes
    "cyl2d_re150_adj": dict(ni=97, nj=193, nk=1, h1=0.002, r_out=30.0, re=150.0, dt=0.005,
                            sigma=1e-3),
}


# the ensemble of 2D cases: circumferential x radial node counts and the Reynolds range
ENSEMBLE_NJ = (129, 193, 257, 385, 513)
ENSEMBLE_NI = (65, 97, 129, 193)
ENSEMBLE_RE = (40.0, 300.0)


def ensemble_cases(n: int = 96, seed: int = SEED, h1: float = 0.002, r_out: float = 30.0,
                   sigma: float = 1e-3, dt: float = 0.005, dt_rule: str = "courant"):
    """Seeded 2D cylinder cases of uneven size and Reynolds number: per case index c, one
    Philox stream (key seed, stream 1000 + c) draws NJ, NI from the size sets and Re uniform in
    [40, 300].  Stretching law and velocity noise as the single 2D case (wall spacing h1, outer
    radius r_out, N(0, sigma) noise with the case's own stream).  dt_rule "courant": the time
    step keeps the 257-node case's circumferential Courant number, dt (256 / (NJ - 1)) capped at
    dt; "uniform": dt for every case."""
    out = []
    for c in range(n):
        gen = np.random.Generator(np.random.Philox(np.random.SeedSequence([seed, 1000 + c])))
        nj = int(ENSEMBLE_NJ[gen.integers(len(ENSEMBLE_NJ))])
        ni = int(ENSEMBLE_NI[gen.integers(len(ENSEMBLE_NI))])
        re = float(gen.uniform(*ENSEMBLE_RE))
        (xg, yg, _), ic = make_case(ni, nj, 1, h1, r_out, sigma, seed=seed + 1 + c)
        cdt = dt if dt_rule == "uniform" else dt * min(1.0, 256.0 / (nj - 1))
        out.append(dict(ni=ni, nj=nj, re=re, dt=cdt, xg=xg, yg=yg, ic=ic, draw=c))
    return out


def make_case(ni, nj, nk, h1, r_out, sigma, seed=SEED, lz=math.pi, **_):
    """Grid nodes and seeded initial fields for a workload (3D when nk > 1)."""
    g = OGrid(ni, nj, nk, h1, r_out, lz)
    xg, yg, zg = g.nodes()
    ic = seeded_ic(ni, nj, nk, sigma, seed, three_d=nk > 1)
    return (xg, yg, zg), ic


if __name__ == "__main__":  # pragma: no cover - manual use
    import argparse
    ap = argparse.ArgumentParser(description="write a reference-format grid + restart")
    ap.add_argument("workload", choices=sorted(WORKLOADS))
    ap.add_argument("--out", default=".")
    a = ap.parse_args()
    w = WORKLOADS[a.workload]
    (xg, yg, zg), ic = make_case(**w)
    os.makedirs(a.out, exist_ok=True)
    if w["nk"] > 1:
        write_grid3d(os.path.join(a.out, grid_filename(w["ni"], w["nj"], w["nk"])), xg, yg, zg)
        write_restart3d(os.path.join(a.out, "RESTART000.DAT"), ic, w["re"], w["re"] * 0.71,
                        w["dt"], 0.0)
    else:
        write_grid2d(os.path.join(a.out, grid_filename(w["ni"], w["nj"])), xg, yg)
        write_restart2d(os.path.join(a.out, "RESTART.DAT"), ic, w["re"], 1.0, w["dt"], 0.0)
