"""Write reference_manifest_steps.json: (family, index, t_start, steps, split) of every
trajectory of the reference's packed dataset, from its data/packed/manifest.json (a data file of the
reference repository, read here as JSON; no reference code is run).  The step counts are those of the
reference's euler2d at 512^2 from each family's initial condition over its recorded window, so they pin a
restatement of the families (npde.families.REFERENCE_FAMILIES) run through a solver that is bit-identical
to the reference's.

  python tests/golden/make_manifest_steps.py [/root/reference/data/packed/manifest.json]
"""
import json
import sys
from pathlib import Path

src = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/data/packed/manifest.json")
traj = json.loads(src.read_text())["trajectories"]
assert all(t["solver_n"] == 512 for t in traj)
rows = [[t["family"], t["index"], t["t_start"], t["steps"], t["split"]] for t in traj]
out = Path(__file__).resolve().parent / "reference_manifest_steps.json"
body = ",\n".join(json.dumps(r) for r in rows)
out.write_text('{"source": "data/packed/manifest.json", "solver_n": 512,\n'
               '"columns": ["family", "index", "t_start", "steps", "split"],\n"rows": [\n' + body + "\n]}\n")
print(f"{len(rows)} trajectories -> {out}")
