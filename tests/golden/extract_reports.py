"""Extract the parity anchors from the reference's published result reports.

Reads docs/results/{validation,profiles,geometry-profiles,cancelled-normals,
tree-metadata,thresholds,torus-100k,torus-1m,torus-1m-d9}.json of the
reference checkout (default /root/reference) and writes reference_reports.json
beside this script: per fixture the generator parameters, the input SHA-256,
per-level support/active key counts and SHA-256s, the isovalues (warp-psr's
and PoissonRecon's), cube_to_model and the basis counts; for the topology
fixtures also the complete tree's node and ghost counts per level and the
input status / aggregate counts; for the threshold fixtures the selected
aggregate's splat depth; for the large torus runs the parameters, basis
counts, cube_to_model and isovalues.  Only data is copied; run once, the
output is committed.
"""
import json
import sys
from pathlib import Path

ROOT = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")
OUT = Path(__file__).resolve().parent / "reference_reports.json"
PARAM_KEYS = ("shape", "count", "seed", "depth", "full_depth", "base_depth", "solve_depth", "kernel_depth",
              "scale", "samples_per_node", "cg_accuracy", "base_v_cycles", "weight", "exact", "iters",
              "precision", "input_scale", "translation", "variant")


def entry(name, source, manifest, rec):
    params = {k: manifest["parameters"].get(k) for k in PARAM_KEYS}
    meta = rec.get("metadata", {})
    return {
        "name": name,
        "source": source,
        "params": params,
        "input_sha256": manifest["input_sha256"],
        "tree": rec["tree"]["levels"],
        "iso_value": rec["iso_value"],
        "reference_iso_value": rec["reference_iso_value"],
        "max_abs_value_error": rec.get("max_abs_value_error"),
        "field_tolerances": rec.get("field_tolerances"),
        "cube_to_model": meta.get("cube_to_model"),
        "basis_counts": meta.get("basis_counts"),
    }


def main():
    out = []
    val = json.loads((ROOT / "docs/results/validation.json").read_text())
    for f in val["fixtures"]:
        out.append(entry(f["fixture"], "validation.json", f["manifest"], f["reconstruction"]))
    for fn in ("profiles.json", "geometry-profiles.json"):
        d = json.loads((ROOT / "docs/results" / fn).read_text())
        for p in d["profiles"]:
            out.append(entry(p["fixture"], fn, p["manifest"], p["reconstruction"]))
    cancelled = json.loads((ROOT / "docs/results/cancelled-normals.json").read_text())
    cases = []
    for c in cancelled["profiles"]:
        cases.append({k: c.get(k) for k in ("case", "input_sha256", "parameters", "iso_value", "reference_iso_value",
                                            "normal_average", "tree")})
    topology = []
    meta = json.loads((ROOT / "docs/results/tree-metadata.json").read_text())
    for p in meta["profiles"]:
        e = entry(p["fixture"], "tree-metadata.json", p["manifest"], p["reconstruction"])
        e["complete_tree"] = p["reconstruction"]["complete_tree"]["levels"]
        e["samples"] = p["reconstruction"].get("samples")
        topology.append(e)
    thresholds = []
    th = json.loads((ROOT / "docs/results/thresholds.json").read_text())
    discoveries = {d["precision"]: d for d in th["discoveries"]}
    for p in th["profiles"]:
        e = entry(p["fixture"], "thresholds.json", p["manifest"], p["reconstruction"])
        e["selected_splat"] = p.get("selected_splat")
        disc = discoveries[p["manifest"]["parameters"]["precision"]]
        e["aggregate_index"] = disc["aggregate_index"]
        e["integer_depth"] = disc["integer_depth"]
        thresholds.append(e)
    large = []
    for name in ("torus-100k", "torus-1m", "torus-1m-d9"):
        r = json.loads((ROOT / "docs/results" / f"{name}.json").read_text())
        m = r["metadata"]
        large.append({"name": name, "source": f"{name}.json", "parameters": m["parameters"], "precision": m["precision"],
                      "basis_counts": m["basis_counts"], "cube_to_model": m["cube_to_model"],
                      "iso_value": r["iso_value"], "reference_iso_value": r["reference_iso_value"],
                      "coefficients": r["coefficients"], "max_abs_value_error": r["max_abs_value_error"]})
    OUT.write_text(json.dumps({"fixtures": out, "cancelled_normals": cases, "complete_tree": topology,
                               "thresholds": thresholds, "large_torus": large}, indent=1) + "\n")
    print(f"wrote {OUT} ({len(out)} fixtures, {len(cases)} cancelled-normal cases, {len(topology)} complete trees, "
          f"{len(thresholds)} threshold fixtures, {len(large)} large torus runs)")


if __name__ == "__main__":
    main()
