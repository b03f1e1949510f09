"""Extracts the reference's committed test fixtures (raw float64 .data files) into
tests/golden/fixtures.json.  Run in a container that has the reference mounted at
/root/reference; the JSON is committed so the tests never read the reference at run time.

Sources (Cuda-CMM/tests/...):
  euler2D_one_step/data/reference-results_4_nodes_C256_F512_t1_T1/Monitoring_data/
  euler2D_11b_4nodes_reference_remapping/data/reference-results_4_nodes_C256_F512_t32_T2/Monitoring_data/
  vlasov1D_15_vp_landau_damping/data/reference-results_C256_F512_t1024_T0.2/Monitoring_data/
  euler2D_init_fields/data/reference-results_<kind>_4_nodes_C256_F512_t1_T0/Monitoring_data/
    (kind = none, comp, conv, sample, zoom, particles, forward: the t = 0 outputs of each save mode)
  euler2D_remap/data/reference-results_4_nodes_C256_F512_t32_T2/Monitoring_data/
    (the second half of a run restarted from the map stack the first half saved)
Also the log.txt warm-up lines "IncompErr = ..." (3 significant digits), the saved params.txt of each
results folder and the input parameter file the case is run with.
"""
import glob
import json
import os
import re

import numpy as np

REF = "/root/reference/Cuda-CMM/tests"
CASES = {
    "one_step": "euler2D_one_step/data/reference-results_4_nodes_C256_F512_t1_T1",
    "remap_11b": "euler2D_11b_4nodes_reference_remapping/data/reference-results_4_nodes_C256_F512_t32_T2",
    "vlasov_15": "vlasov1D_15_vp_landau_damping/data/reference-results_C256_F512_t1024_T0.2",
    "restart": "euler2D_remap/data/reference-results_4_nodes_C256_F512_t32_T2",
}
INIT_KINDS = ("none", "comp", "conv", "sample", "zoom", "particles", "forward")
for _k in INIT_KINDS:
    CASES["init_" + _k] = f"euler2D_init_fields/data/reference-results_{_k}_4_nodes_C256_F512_t1_T0"
# the parameter files the reference's harness runs each case with (Cuda-CMM/tests/*/run_test.py)
INPUTS = {
    "one_step": "euler2D_one_step/params.txt",
    "remap_11b": "../examples/euler2D/params-11b-4nodes-reference-remapping.txt",
    "vlasov_15": "vlasov1D_15_vp_landau_damping/params-15b-vp-landau-damping.txt",
    # the restarted half; its first half (saving the map stack) is params_input_first
    "restart": "euler2D_remap/params-2.txt",
}
FIRST_INPUTS = {"restart": "euler2D_remap/params-1.txt"}
for _k in INIT_KINDS:
    INPUTS["init_" + _k] = f"euler2D_init_fields/params_{_k}.txt"


def main():
    out = {}
    for case, rel in CASES.items():
        base = os.path.join(REF, rel)
        d = {}
        for f in sorted(glob.glob(os.path.join(base, "Monitoring_data", "**", "*.data"), recursive=True)):
            if f.endswith("_Hash.data"):
                continue
            key = os.path.relpath(f, os.path.join(base, "Monitoring_data"))[:-5]
            d[key] = np.fromfile(f, dtype="<f8").tolist()
        log = open(os.path.join(base, "log.txt")).read()
        d["log_warmup_incomp"] = [float(x) for x in re.findall(r"IncompErr = ([0-9.e+-]+)", log)]
        # params.txt in the results folder is the reference's own save_param_file output
        d["params_saved"] = open(os.path.join(base, "params.txt")).read()
        d["params_input"] = open(os.path.join(REF, INPUTS[case])).read()
        if case in FIRST_INPUTS:
            d["params_input_first"] = open(os.path.join(REF, FIRST_INPUTS[case])).read()
        out[case] = d
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=0, sort_keys=True)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
