"""The Fortran driver (fvm_amd/bin/fvm_rles) runs the reference's own input files -- the grid
file NNNxNNNxNNN.dat and RESTART000.DAT -- through the C ABI and writes the reference's restart,
NU_CD_CL.PLT, CS.PLT and CT.PLT; its result after two converged steps, the passive temperature
included, matches the reference's golden dump."""
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO, load_golden

pytestmark = pytest.mark.gpu

EXE = os.path.join(REPO, "fvm_amd", "bin", "fvm_rles")


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built (flang absent)")
def test_fortran_driver_two_steps(tmp_path):
    from fvm_amd.grid import grid_filename, read_restart3d, write_grid3d, write_restart3d
    g = load_golden("rles_9x21x5_les_thermal")
    ni, nj, nk = int(g["ni"]), int(g["nj"]), int(g["nk"])
    write_grid3d(str(tmp_path / grid_filename(ni, nj, nk)), g["xg"], g["yg"], g["zg"])
    f = {k: g["in_" + k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz", "t", "adv")}
    write_restart3d(str(tmp_path / "RESTART000.DAT"), f, float(g["re"]), 0.0, float(g["dt"]), 0.0)
    (tmp_path / "fvm_rles.nml").write_text(
        f"&FVM_RUN NI={ni}, NJ={nj}, NK={nk}, RE={float(g['re'])!r}, DTIME={float(g['dt'])!r},\n"
        f" RESIMAX=1.0D-13, NSWPP=50, USE_LES=1, REF_A={float(g['ref_a'])!r},"
        f" POISSON_MODE=2, NSTEP=2, THERMAL=1, PE={float(g['pe'])!r}, NITER_W=2 /\n")
    print("fortran driver: rles_9x21x5, 2 steps", flush=True)
    out = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    r = read_restart3d(str(tmp_path / "RESTART000.DAT"), ni, nj, nk)
    for c in ("u", "v", "w"):
        assert rel(r[c], g[f"s2_{c}"]) < 1e-10, c
    assert rel(r["p"], g["s2_pref"]) < 1e-10
    assert rel(r["t"], g["s2_t"]) < 1e-10
    assert r["niter"] == 2 and abs(r["rtime"] - 2 * float(g["dt"])) < 1e-15
    body = (tmp_path / "NU_CD_CL.PLT").read_text().split("\n", 1)[1]
    rows = np.array(body.split(), dtype=np.float64).reshape(-1, 10)
    assert rows.shape == (2, 10)
    for st in (0, 1):
        ref = g["scalars"][st]
        assert abs(rows[st, 2] - ref[6]) <= 1e-9 * abs(ref[6])      # CD
        assert abs(rows[st, 3] - ref[7]) <= 1e-9 * max(abs(ref[7]), 1.0)  # CL
        assert abs(rows[st, 9] - ref[4]) <= 1e-10 * ref[4]           # CFL
        assert abs(rows[st, 1] - ref[10]) <= 1e-9 * abs(ref[10])     # NU
    # nodal Tecplot fields of the last step (D16.9 records: 10 significant digits)
    nn = ni * nj * nk
    for fname, cols in (("PVT_001_000.PLT", ("out_pg", "out_ug", "out_vg", "out_wg", "out_tg")),
                        ("VOR_001_000.PLT", ("out_dum3d", "out_lamdag"))):
        lines = (tmp_path / fname).read_text().splitlines()
        assert lines[2].split(",")[1].split("=")[1].strip() == str(ni)
        vals = np.array(" ".join(lines[3:]).replace("D", "e").split(), dtype=np.float64)
        vals = vals.reshape(nn, 3 + len(cols))
        for q, key in enumerate(cols):
            ref = g[key][1:, 1:, 1:] if key == "out_dum3d" else g[key]
            ref = ref.ravel(order="F")
            assert np.abs(vals[:, 3 + q] - ref).max() <= 1e-8 * np.abs(ref).max(), (fname, key)
    # wall y+ contour of the last step on the wall cells, with their centres
    lines = (tmp_path / "Ypl_001_000.PLT").read_text().splitlines()
    assert [int(f.split("=")[1]) for f in lines[2].split(",")[1:4]] == [1, nj - 1, nk - 1]
    vals = np.array(" ".join(lines[3:]).replace("D", "e").split(), dtype=np.float64)
    vals = vals.reshape((nj - 1) * (nk - 1), 4)
    for q, ref in enumerate((g["setup_x"][1, 1:nj, 1:nk], g["setup_y"][1, 1:nj, 1:nk],
                             g["setup_z"][1, 1:nj, 1:nk], g["out_yplus"])):
        ref = ref.ravel(order="F")
        assert np.abs(vals[:, q] - ref).max() <= 1e-8 * np.abs(ref).max(), ("Ypl", q)
    for fname, key in (("CS.PLT", "cs_sqrt"), ("CT.PLT", "ct_sqrt")):
        body = (tmp_path / fname).read_text().split("\n", 1)[1]
        vals = np.array(body.split(), dtype=np.float64).reshape(-1, 2)
        assert np.abs(vals[:, 1] - g[key]).max() <= 1e-10 * np.abs(g[key]).max(), fname


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built (flang absent)")
def test_fortran_driver_post_process(tmp_path):
    """One step to RTIME_MAX, then POST_PROCESS's averaging phase at DTIME_POST up to
    RTIME_END (three steps) and its Avrg000 / Urms000 / Strs000 / Surf000.PLT files, against
    the values and headers the reference wrote for the same run."""
    from fvm_amd.grid import grid_filename, write_grid3d, write_restart3d
    g = load_golden("rles_9x21x5_les_post")
    ni, nj, nk = int(g["ni"]), int(g["nj"]), int(g["nk"])
    write_grid3d(str(tmp_path / grid_filename(ni, nj, nk)), g["xg"], g["yg"], g["zg"])
    f = {k: g["in_" + k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz", "t", "adv")}
    write_restart3d(str(tmp_path / "RESTART000.DAT"), f, float(g["re"]), 0.0, float(g["dt"]), 0.0)
    (tmp_path / "fvm_rles.nml").write_text(
        f"&FVM_RUN NI={ni}, NJ={nj}, NK={nk}, RE={float(g['re'])!r}, DTIME={float(g['dt'])!r},\n"
        f" RESIMAX=1.0D-13, NSWPP=50, USE_LES=1, REF_A={float(g['ref_a'])!r},"
        f" POISSON_MODE=2, THERMAL=1, PE={float(g['pe'])!r}, RTIME_MAX={float(g['dt'])!r},"
        f" RTIME_END=0.0065, DTIME_POST={float(g['dt_post'])!r} /\n")
    print("fortran driver: rles_9x21x5, 1 step + 3 averaging steps", flush=True)
    out = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    body = (tmp_path / "NU_CD_CL.PLT").read_text().split("\n", 1)[1]
    assert np.array(body.split(), dtype=np.float64).reshape(-1, 10).shape == (4, 10)
    headers = g["plt_headers"]
    for h, (fname, cols, nrec) in enumerate((
            ("Avrg000.PLT", ("pg", "ug", "vg", "wg", "tg"), ni * nj * nk),
            ("Urms000.PLT", ("urms", "vrms", "wrms"), ni * nj * nk),
            ("Strs000.PLT", ("uu", "uv", "uw", "vv", "vw", "ww"), ni * nj * nk),
            ("Surf000.PLT", ("surf_theta", "surf_cf", "surf_cp", "surf_nu"),
             (nj - 1) * (nk - 1)))):
        lines = (tmp_path / fname).read_text().splitlines()
        assert lines[:3] == list(headers[h]), fname
        vals = np.array(" ".join(lines[3:]).replace("D", "e").split(), dtype=np.float64)
        vals = vals.reshape(nrec, 3 + len(cols))
        for q, key in enumerate(cols):
            ref = g["plt_" + key].ravel(order="F")
            assert np.abs(vals[:, 3 + q] - ref).max() <= 1e-8 * np.abs(ref).max(), (fname, key)


EXE2D = os.path.join(REPO, "fvm_amd", "bin", "fvm_r2d")


@pytest.mark.skipif(not os.path.exists(EXE2D), reason="Fortran driver not built (flang absent)")
def test_fortran_2d_driver_three_steps(tmp_path):
    """The 2D driver reads the 2D reference's NNNxNNN.dat and list-directed RESTART.DAT and
    reproduces its three converged steps."""
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
    print("fortran 2D driver: r2d_17x33, 3 steps", flush=True)
    out = subprocess.run([EXE2D], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    r = read_restart2d(str(tmp_path / "RESTART.DAT"), ni, nj)
    for c in ("u", "v", "p"):
        assert rel(r[c][:, :, 0], g[f"s3_{c}"]) < 1e-10, c
    body = (tmp_path / "NU_CD_CL.PLT").read_text().split("\n", 1)[1]
    rows = np.array(body.split(), dtype=np.float64).reshape(-1, 10)
    assert rows.shape == (3, 10)
    for st in range(3):
        ref = g["scalars"][st]  # step niter resip cfl cpb cd cl cd_p cd_v cl_p cl_v
        assert abs(rows[st, 2] - ref[5]) <= 1e-9 * abs(ref[5])             # CD
        assert abs(rows[st, 3] - ref[6]) <= 1e-9 * max(abs(ref[6]), 1.0)   # CL
