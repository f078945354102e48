"""Writes tests/parity_summary.md from the JSON-lines records that the GPU
parity tests append to $SLRH_PARITY_OUT:

    SLRH_PARITY_OUT=/tmp/parity.jsonl python -u -m pytest tests -m gpu -x -v -s
    python tests/parity_report.py /tmp/parity.jsonl
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(path):
    recs = [json.loads(l) for l in open(path) if l.strip()]
    out = ["# Parity summary: GPU path against the reference and the oracle", "",
           "Written by `tests/parity_report.py` from the `SLRH_PARITY_OUT` records of one",
           "`python -m pytest tests -m gpu` run on an MI355X.", "",
           "## Images against the compiled reference renderer (3 sigma per pixel and CIE XYZ channel)", "",
           "Gate (`tests/refcases.py` `check_3sigma`): at most 1 % of pixel-channels outside 3 combined standard",
           "errors, at most max(6, 4 %) of any 8x8 tile's entries outside (about 1 expected by chance), image total",
           "within 4 standard errors, relMSE < 0.05.", "",
           "| case | pixels | batches (GPU, reference) | spp per reference batch | outside 3 sigma | worst 8x8 tile |"
           " relMSE | total z |", "|---|---|---|---|---|---|---|---|"]
    for r in recs:
        if "fail_frac" in r:
            out.append("| %s | %d | %d, %d | %d | %.2f %% | %d | %.2e | %.2f |" % (
                r["case"], r["pixels"], r["batches"][0], r["batches"][1], r["spp_ref"], 100 * r["fail_frac"],
                r["worst_tile"], r["relmse"], r["total_z"]))
    out += ["", "## Same-seed films against the oracle (per-material shading on a smooth-normal sphere)", "",
            "Gate (`tests/test_gpu_bsdf.py`): more than 97 % of pixel strata within 1e-4 relative, image totals",
            "within 2e-3.", "", "| case | strata within 1e-4 | total relative difference | rays relative difference |",
            "|---|---|---|---|"]
    for r in recs:
        if "strata_within_1e-4" in r:
            out.append("| %s | %.3f %% | %.1e | %.1e |" % (r["case"], 100 * r["strata_within_1e-4"], r["total_rel"],
                                                         r["rays_rel"]))
    out += ["", "## Ray batches against the reference QBVH", "",
            "`tests/test_gpu_c2.py::test_c2_ray_batch_parity`: 2^20 rays through the C2 scene (1,310,756",
            "triangles): hit triangle, t and surface parameters bit-identical to `QBVH::intersect` (ties within",
            "1e-6 in t excepted; none occur), visibility identical to `QBVH::testVisibility`.", ""]
    with open(os.path.join(HERE, "parity_summary.md"), "w") as f:
        f.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main(sys.argv[1])
