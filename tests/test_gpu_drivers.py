"""The command-line drivers (`hydro`, `hydro3d`) against the reference's step loop run
on the CPU oracle: the same stdout lines (initial mass and Dt, per-iteration t and Dt,
mass lines) and byte-identical CSV files, for the reference's own small configuration
(hydro2d README example: 60 x 60, 200 iterations, t_max 2, write every 10) and for the
added boundary-mode and precision options."""
import pathlib
import re
import subprocess

import numpy as np
import pytest

import reference_loop

pytestmark = pytest.mark.gpu

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG = ROOT / "scalable-fpga-hydrodynamics_amd"

KEEP = re.compile(r"^(Initial total mass .*|Initial Dt .*|\[it=\s*\d+\] t=.*|\[it=\s*\d+\] hydro_fvm compute time "
                  r"t=\S+ \(Dt=\S+\))")


def _driver_lines(stdout: str):
    out = []
    for line in stdout.splitlines():
        m = KEEP.match(line)
        if m:
            out.append(m.group(1))
    return out


def _run(binary, args, cwd):
    exe = PKG / binary
    assert exe.exists(), f"{exe} is not built"
    p = subprocess.run([str(exe), *args], cwd=cwd, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    return p.stdout


CASES = [
    ("hydro", (60, 60), 2.0, 200, 10, ["-j", "2", "-k", "1"], np.float64, None),
    ("hydro", (60, 60), 2.0, 200, 10, ["--sync-steps"], np.float64, None),
    # no CSV files, but the same mass lines (the host long double sum of hydro2d/main.cxx:237-265)
    ("hydro", (60, 60), 2.0, 200, 10, ["--no-output"], np.float64, None),
    ("hydro3d", (24, 20, 16), 2.0, 40, 15, ["--no-output", "--sync-steps"], np.float64, None),
    ("hydro", (50, 40), 2.0, 60, 20, ["--precision", "fp32"], np.float32, None),
    ("hydro", (40, 30), 2.0, 50, 25, ["--bcx", "reflective", "--bcy", "periodic"], np.float64, (1, 0)),
    ("hydro", (70, 33), 0.01, 1000, 5, [], np.float64, None),
    ("hydro3d", (24, 20, 16), 2.0, 40, 15, [], np.float64, None),
    ("hydro3d", (16, 12, 20), 2.0, 30, 10, ["--sync-steps", "--bcz", "periodic"], np.float64, (0, 1, 0)),
    ("hydro3d", (22, 18, 26), 2.0, 30, 10, ["--precision", "fp32"], np.float32, None),
    # Several 131072-row blocks per file (parallel formatting, in-order writes) and a
    # file written in the background while the next steps run.
    ("hydro", (700, 400), 2.0, 12, 5, [], np.float64, None),
    ("hydro3d", (60, 50, 48), 2.0, 6, 4, [], np.float64, None),
    # The reference on MPI ranks: CSV rows rank by rank, each rank's subdomains in turn,
    # and the mass summed per rank (the first is its own small configuration, mpirun -np 2
    # --sdx 2 --sdy 1).
    ("hydro", (60, 60), 2.0, 200, 10, ["-j", "2", "-k", "1", "--ranks", "2"], np.float64, None),
    ("hydro", (50, 40), 2.0, 60, 20, ["-j", "3", "--ranks", "2"], np.float64, None),
    ("hydro", (41, 30), 2.0, 30, 10, ["-j", "5", "--ranks", "3"], np.float64, None),
    ("hydro3d", (21, 18, 16), 2.0, 20, 10, ["-j", "4", "--ranks", "3"], np.float64, None),
]


def _split(extra, dim):
    """(sd, ranks) of the driver arguments, or (None, 1) for the whole domain's order
    (one rank with an X-only split walks the cells in that order)."""
    sd = [2, 1, 1][:dim]
    ranks = 1
    for flag, val in zip(extra, extra[1:] + [""]):
        if flag in ("-j", "-k", "-l"):
            sd["jkl".index(flag[1])] = int(val)
        if flag == "--ranks":
            ranks = int(val)
    if ranks == 1 and all(s == 1 for s in sd[1:]):
        return None, 1
    return tuple(sd), ranks


@pytest.mark.parametrize("binary,nb,max_t,max_it,w,extra,dtype,bc", CASES)
def test_driver_matches_reference_loop(gpu, oracle, tmp_path, binary, nb, max_t, max_it, w, extra, dtype, bc):
    args = ["-x", str(nb[0]), "-y", str(nb[1])] + (["-z", str(nb[2])] if len(nb) == 3 else [])
    args += ["-t", str(max_t), "-i", str(max_it), "-w", str(w), "-o", "gpu"] + extra
    stdout = _run(binary, args, tmp_path)
    ref_dir = tmp_path / "ref"
    ref_dir.mkdir()
    sd, ranks = _split(extra, len(nb))
    want_lines, want_files = reference_loop.run(oracle, nb, max_t, max_it, w, ref_dir / "gpu", dtype=dtype, bc=bc,
                                                sd=sd, ranks=ranks)
    got_lines = _driver_lines(stdout)
    assert got_lines == want_lines
    if "--no-output" in extra:
        want_files = []
    got_files = sorted(p.name for p in tmp_path.glob("gpu_*.csv"))
    assert got_files == sorted(p.name for p in want_files)
    for f in want_files:
        assert (tmp_path / f.name).read_bytes() == f.read_bytes(), f.name


def test_driver_refuses_y_and_z_splits(gpu, tmp_path):
    """The reference's kernel makes every subdomain's own Y and Z edges walls
    (hydro2d/kernel.cxx:214, 235), so a Y or Z split is a different problem; the drivers
    reproduce X splits and refuse the others."""
    for exe, args in (("hydro", ["-k", "2"]), ("hydro3d", ["-l", "2"]), ("hydro3d", ["-k", "3"])):
        p = subprocess.run([str(PKG / exe), "-x", "8", "-y", "8", *(["-z", "8"] if exe == "hydro3d" else []),
                            "-i", "2", "--no-output", *args], capture_output=True, text=True, timeout=60, cwd=tmp_path)
        assert p.returncode != 0 and "split along X only" in p.stderr, (exe, args, p.stderr)


def test_driver_help_and_bad_option(gpu, tmp_path):
    exe = PKG / "hydro"
    p = subprocess.run([str(exe), "-h"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0 and "Hydrodynamics 2D" in p.stdout and "--bcx" in p.stdout
    p = subprocess.run([str(exe), "--bcy", "sideways"], capture_output=True, text=True, timeout=60)
    assert p.returncode != 0 and "Unknown boundary mode" in p.stderr


def test_driver_refuses_sizes_beyond_the_device(gpu, tmp_path):
    """Sizes are read as the reference reads them (size_t(stoi(...))), so -x -5 becomes a
    huge size; the drivers' memory check, done in floating point, refuses it before
    allocating anything (a size_t product would wrap to a small footprint)."""
    for exe, args in (("hydro", ["-x", "-5", "-y", "10"]), ("hydro3d", ["-x", "10", "-y", "-3", "-z", "10"]),
                      ("hydro3d", ["-x", "2000000", "-y", "2000000", "-z", "2000000"])):
        p = subprocess.run([str(PKG / exe), *args, "-i", "1", "--no-output"], capture_output=True, text=True,
                           timeout=60, cwd=tmp_path)
        assert p.returncode != 0 and "Insufficient device memory" in p.stderr, (exe, args, p.stderr[-400:])
