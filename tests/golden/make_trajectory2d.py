"""Write the 2D trajectory fixture (runs in the build container only).

A developed, shedding state of the 2D trajectory case (65 radial x 129 circumferential nodes,
wall spacing 0.002 D, outer radius 30 D, Re 100, dt 0.0025), given as an .npz of the fields u, v,
p, nlx, nly (and the grid nodes xg, yg, RTIME), is written as the reference's own input files
(NNNxNNN.dat, RESTART.DAT); the reference 2D solver (flang build of
cylinder/cylinder_v3.0.3_AF_original 4step_face_nl by oracle/ref/build_ref2d.sh, trajectory
harness oracle/ref/harness2d_post.f90, Poisson line Gauss-Seidel to RESIMAX 1e-8) then runs NSTEP
steps from it in its post-processing phase.  Kept: the inputs as the reference parsed them, its
per-step C_D, C_L and Poisson sweeps / RESIP, the last row of POST_NU_CD_CL.PLT and
TIME_AVERAGED_RESULTS.DAT.

usage: python tests/golden/make_trajectory2d.py STATE.npz NSTEP
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
from fvm_amd.grid import grid_filename, read_restart2d, write_grid2d, write_restart2d  # noqa: E402

LABEL = "r2d_65x129_re100_shedding"


def main(state_path: str, nstep: int) -> None:
    st = np.load(state_path)
    xg, yg = st["xg"], st["yg"]
    ni, nj = xg.shape
    re, dt, rtime = float(st["re"]), float(st["dt"]), float(st["rtime"])
    exe = os.path.join(REPO, "oracle", "_ref", "r2d_65x129_traj", "harness2d_post")
    if not os.path.exists(exe):
        subprocess.run(["bash", os.path.join(REPO, "oracle", "ref", "build_ref2d.sh"),
                        "r2d_65x129_traj", str(ni), str(nj), "2000000", "1.0D-8",
                        "harness2d_post"], check=True)
    out = {"ni": ni, "nj": nj, "re": re, "dt": dt, "rtime0": rtime, "nstep": nstep,
           "resimax": 1e-8, "xg": xg, "yg": yg}
    with tempfile.TemporaryDirectory() as d:
        write_grid2d(os.path.join(d, grid_filename(ni, nj)), xg[:, :, None], yg[:, :, None])
        z = np.zeros((ni + 1, nj + 1, 1))
        f = {k: np.asarray(st[k])[:, :, None] for k in ("u", "v", "p", "nlx", "nly")}
        f.update(adv=z, t=z)
        write_restart2d(os.path.join(d, "RESTART.DAT"), f, re, 0.71, dt, 0.0, rtime=rtime)
        rs = read_restart2d(os.path.join(d, "RESTART.DAT"), ni, nj)
        for k in ("u", "v", "p", "nlx", "nly"):
            out["in_" + k] = rs[k][:, :, 0]
        subprocess.run([exe, str(nstep)], cwd=d, check=True, stdout=subprocess.DEVNULL,
                       input=("\n" * 100000).encode())
        rows = [l.split() for l in open(os.path.join(d, "scalars.txt")) if l.startswith("step")]
        # columns: step, niter, resip, cfl, cpb, cd, cl, cd_p, cd_v, cl_p, cl_v
        out["scalars"] = np.array([[float(x) for x in r[1:]] for r in rows])
        # list-directed records (a record may span several lines): every number after the
        # header line, nine per record
        num = lambda path: np.array([float(x.replace("D", "e")) for x in
                                     open(path).read().split("\n", 1)[1].split()])
        # TIME NU CD CL CL_AMP CPB ST RMS_CD RMS_CL per step
        out["post"] = num(os.path.join(d, "POST_NU_CD_CL.PLT")).reshape(-1, 9)
        out["final"] = num(os.path.join(d, "TIME_AVERAGED_RESULTS.DAT"))[:9]
    np.savez_compressed(os.path.join(HERE, LABEL + ".npz"), **out)
    print("wrote", LABEL, "steps", nstep, "final", out["final"])


if __name__ == "__main__":
    main(sys.argv[1], int(sys.argv[2]))
