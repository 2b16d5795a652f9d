"""Extract the numeric pins held in the reference's published run logs.

Reads hydro2d/reference_run/res_*.res and hydro3d/reference_run/res_*.res under
/root/reference (this script runs only where the reference is mounted) and writes
tests/golden/res_logs.json: for each run its grid, rank count, limits, the printed
initial total mass and initial Dt, every per-iteration "(t, Dt)" pair and the final
mass line.  These are data (the logs' printed values), not reference source.

Line formats (hydro2d/main.cxx:673-674, 809-811, 273-274):
  "Initial total mass       : %.7e"
  "Initial Dt : %.2e"
  "[it=%4lu] hydro_fvm compute time t=%.4e (Dt=%.3e) : ..."
  "[it=%4lu] t=%.4e Total density : %.7e, change in mass: %.1e ..."
  "Mass change : %.2e"                                  (hydro2d/main.cxx:887)
"""
import json
import pathlib
import re
import sys

REF = pathlib.Path("/root/reference")
OUT = pathlib.Path(__file__).with_name("res_logs.json")

RUN_RE = re.compile(r"res_(\d+)_(\d+)_(\d+)(?:_(\d+))?(_2FPGAS)?\.res$")
IT_RE = re.compile(r"^\[it=\s*(\d+)\] hydro_fvm compute time t=(\S+) \(Dt=(\S+)\)")
MASS_RE = re.compile(r"^\[it=\s*(\d+)\] t=(\S+) Total density : (\S+), change in mass: (\S+)")


def parse(path: pathlib.Path, dim: int) -> dict:
    text = path.read_text().splitlines()
    cfg = {}
    for line in text:
        for key, lab in (("nb_x", "Spatial Points X axis"), ("nb_y", "Spatial Points Y axis"),
                         ("nb_z", "Spatial Points Z axis"), ("max_t", "Max Timestep (max_t)"),
                         ("max_it", "Max iterations (max_it)")):
            if lab in line and key not in cfg:
                cfg[key] = line.split(":")[-1].strip()
    traj = []
    masses = []
    init_mass = init_dt = mass_change = None
    for line in text:
        m = IT_RE.match(line)
        if m:
            traj.append([int(m.group(1)), m.group(2), m.group(3)])
            continue
        m = MASS_RE.match(line)
        if m:
            masses.append([int(m.group(1)), m.group(2), m.group(3), m.group(4)])
            continue
        if line.startswith("Initial total mass"):
            init_mass = line.split(":")[-1].strip()
        elif line.startswith("Initial Dt"):
            init_dt = line.split(":")[-1].strip()
        elif line.startswith("Mass change :"):
            mass_change = line.split(":")[-1].strip()
    for i, row in enumerate(traj):
        assert row[0] == i, (path, i, row)
    return {
        "file": str(path.relative_to(REF)),
        "dim": dim,
        "nb": [int(cfg["nb_x"]), int(cfg["nb_y"])] + ([int(cfg["nb_z"])] if dim == 3 else []),
        "ranks": 2 if "2FPGAS" in path.name else 1,
        "max_t": float(cfg["max_t"]),
        "max_it": int(cfg["max_it"]),
        "initial_mass": init_mass,
        "initial_dt": init_dt,
        "trajectory": traj,
        "mass_lines": masses,
        "mass_change": mass_change,
    }


def main() -> int:
    if not REF.is_dir():
        print("reference not mounted; nothing to extract", file=sys.stderr)
        return 1
    runs = {}
    for dim, sub in ((2, "hydro2d"), (3, "hydro3d")):
        for p in sorted((REF / sub / "reference_run").glob("res_*.res")):
            key = f"{dim}d_" + "x".join(p.stem.split("_")[2:])
            runs[key] = parse(p, dim)
    OUT.write_text(json.dumps(runs, separators=(",", ":")))
    for k, v in runs.items():
        print(k, v["nb"], v["ranks"], v["initial_mass"], v["initial_dt"], len(v["trajectory"]), v["mass_lines"][-1], v["mass_change"])
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
