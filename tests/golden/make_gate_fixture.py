"""Write tests/golden/gate_reference.json: coarse gates the reference itself
published, with the profile recipe and budget each was computed for.

Sources (read as data; nothing of the reference is imported or run):
  * docs/audits/gate-model-2026-09-26/gpu-band-sweep.json -- field
    ``float64_reference_gate`` per (fd, band) at n=4096, snr 5.5; the gates
    come from the reference gate model on the teaser reference profile
    (tools/audit_gpu_bands.py -> tools/teaser_figure.py:_reference ->
    tests/test_api.py:inspiral_power(4096)).
  * docs/audits/gate-model-2026-09-26/fdr-1e-4.json -- field ``gate`` at
    n=1024, band 256, snr 5.5, fd 1e-4, on the benchmark profile
    (tools/audit_gate_model.py -> matchedfilter.benchmark._inspiral_power).

Run: python tests/golden/make_gate_fixture.py [reference root]
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
AUDIT = os.path.join("docs", "audits", "gate-model-2026-09-26")


def main(root="/root/reference"):
    with open(os.path.join(root, AUDIT, "gpu-band-sweep.json")) as fh:
        sweep = json.load(fh)
    with open(os.path.join(root, AUDIT, "fdr-1e-4.json")) as fh:
        audit = json.load(fh)
    cases = [{"profile": "test_api_inspiral", "n": sweep["n"], "band": r["band"], "snr": sweep["snr"],
              "fd": r["fd"], "gate": r["float64_reference_gate"],
              "source": AUDIT + "/gpu-band-sweep.json"} for r in sweep["rows"]]
    cases.append({"profile": "benchmark_inspiral", "n": audit["n"], "band": audit["band"],
                  "snr": audit["snr"], "fd": audit["fd"], "gate": audit["gate"],
                  "source": AUDIT + "/fdr-1e-4.json"})
    with open(os.path.join(HERE, "gate_reference.json"), "w") as fh:
        json.dump({"cases": cases}, fh, indent=1)
        fh.write("\n")


if __name__ == "__main__":
    main(*sys.argv[1:])
