"""The reference's MAIN written out statement for statement against the reference's own
subroutine interfaces -- ITER_U(IN,INM,INA,INB,INC,IND), NL(IN,OUT), DIF(IN,OUT), BCU(A,B),
P_SOURCE(P), BCP, ... (LES/v2.3_LES-Reynolds stresses/main.f90:35-161) and the 2D ITER_U / ITER_V /
BCU / BCV (cylinder/.../Main.f90:48-118) -- linked against the Fortran shims that bind those
names to the per-subroutine C entry points (fvm_amd/bin/fvm_rles_shim, fvm_r2d_shim).  Their
converged steps reproduce the reference's golden dumps to 1e-10, as the one-call-per-phase drivers
do."""
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO, load_golden

pytestmark = pytest.mark.gpu

EXE3 = os.path.join(REPO, "fvm_amd", "bin", "fvm_rles_shim")
EXE2 = os.path.join(REPO, "fvm_amd", "bin", "fvm_r2d_shim")


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.skipif(not os.path.exists(EXE3), reason="Fortran shim driver not built (flang absent)")
def test_rles_main_through_reference_interfaces(tmp_path):
    from fvm_amd.grid import grid_filename, read_restart3d, write_grid3d, write_restart3d
    g = load_golden("rles_9x21x5_les_converged")
    ni, nj, nk = int(g["ni"]), int(g["nj"]), int(g["nk"])
    write_grid3d(str(tmp_path / grid_filename(ni, nj, nk)), g["xg"], g["yg"], g["zg"])
    z = np.zeros_like(g["in_u"])
    f = {k: g["in_" + k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz")}
    f.update(t=z, adv=z)
    write_restart3d(str(tmp_path / "RESTART000.DAT"), f, float(g["re"]), 0.0, float(g["dt"]), 0.0)
    (tmp_path / "fvm_rles.nml").write_text(
        f"&FVM_RUN NI={ni}, NJ={nj}, NK={nk}, RE={float(g['re'])!r}, DTIME={float(g['dt'])!r},\n"
        f" RESIMAX=1.0D-13, NSWPP=50, USE_LES=1, REF_A={float(g['ref_a'])!r},"
        f" POISSON_MODE=2, NSTEP=2 /\n")
    print("reference MAIN through the shims: rles_9x21x5, 2 steps", flush=True)
    out = subprocess.run([EXE3], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    r = read_restart3d(str(tmp_path / "RESTART000.DAT"), ni, nj, nk)
    for c in ("u", "v", "w"):
        assert rel(r[c], g[f"s2_{c}"]) < 1e-10, c
    assert rel(r["p"], g["s2_pref"]) < 1e-10
    body = (tmp_path / "NU_CD_CL.PLT").read_text().split("\n", 1)[1]
    rows = np.array(body.split(), dtype=np.float64).reshape(-1, 10)
    for st in (0, 1):
        ref = g["scalars"][st]  # step niter resip pref cfl cpb cd cl cd_p cd_v
        assert abs(rows[st, 2] - ref[6]) <= 1e-9 * abs(ref[6])
        assert abs(rows[st, 3] - ref[7]) <= 1e-9 * max(abs(ref[7]), 1.0)
        assert abs(rows[st, 9] - ref[4]) <= 1e-10 * ref[4]


@pytest.mark.skipif(not os.path.exists(EXE2), reason="Fortran shim driver not built (flang absent)")
def test_r2d_main_through_reference_interfaces(tmp_path):
    from fvm_amd.grid import grid_filename, read_restart2d, write_grid2d, write_restart2d
    g = load_golden("r2d_17x33_re100_converged")
    ni, nj = int(g["ni"]), int(g["nj"])
    write_grid2d(str(tmp_path / grid_filename(ni, nj)), g["xg"][:, :, None], g["yg"][:, :, None])
    z = np.zeros((ni + 1, nj + 1, 1))
    f = {k: g["in_" + k][:, :, None] for k in ("u", "v", "p", "nlx", "nly")}
    f.update(adv=z, t=z)
    write_restart2d(str(tmp_path / "RESTART.DAT"), f, float(g["re"]), 0.71, float(g["dt"]), 0.0)
    (tmp_path / "fvm_r2d.nml").write_text(
        f"&FVM_RUN NI={ni}, NJ={nj}, RESIMAX=1.0D-15, NSWPP=50, POISSON_MODE=2, NSTEP=3 /\n")
    print("reference 2D MAIN through the shims: r2d_17x33, 3 steps", flush=True)
    out = subprocess.run([EXE2], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    r = read_restart2d(str(tmp_path / "RESTART.DAT"), ni, nj)
    for c in ("u", "v", "p"):
        assert rel(r[c][:, :, 0], g[f"s3_{c}"]) < 1e-10, c
    body = (tmp_path / "NU_CD_CL.PLT").read_text().split("\n", 1)[1]
    rows = np.array(body.split(), dtype=np.float64).reshape(-1, 10)
    for st in range(3):
        ref = g["scalars"][st]  # step niter resip cfl cpb cd cl cd_p cd_v cl_p cl_v
        assert abs(rows[st, 2] - ref[5]) <= 1e-9 * max(abs(ref[5]), 1.0)
        assert abs(rows[st, 3] - ref[6]) <= 1e-9 * max(abs(ref[6]), 1.0)
