"""Extracts golden vectors from the reference's own test fixtures into tests/golden/.

Run in the build container only (it reads the read-only reference checkout); the outputs are
small committed data files, so nothing at run time reads the reference.

Sources (fixtures of crates/fereg-core/tests, written there by fixest 0.14.2):
  tests/data/fit_d1.csv, fit_toy.csv          the input panels
  tests/data/fit_fixest.json                  OLS cases (tests/fit_ols.rs:10-67, RTOL 1e-8)
  tests/data/fit_fixest_glm.json              Poisson cases (tests/fit_glm.rs)
"""
import json
import os
import shutil

SRC = "/root/reference/crates/fereg-core/tests/data"
DST = os.path.dirname(os.path.abspath(__file__))

OLS_CASES = ["fe2_iid", "fe2_hetero", "fe2_cl_g", "fe2_cl_c1", "fe2_cl_c3", "fe2_iid_ssc_none",
             "fe2_iid_ssc_full", "fe2_iid_ssc_noadj", "fe2_c1_c3_ssc_nog", "fe3_iid_exact",
             "fe1_w_iid", "fe1_w_hetero", "fe1_w_cl", "fe_g_c1_cl_c1", "collin",
             "toy_iid", "toy_hetero", "toy_cl_g", "toy_cl_g_full", "toy_cl_g_none"]
POIS_CASES = ["pois_fe2_iid", "pois_fe2_hetero", "pois_collin", "pois_inter_fx", "pois_fe2_off_w"]
KEEP = ["coef", "se", "vcov", "df_k", "nobs", "nparams", "collin", "obs", "G", "resid", "ssr", "coef_names"]


def main():
    for f in ["fit_d1.csv", "fit_toy.csv"]:
        shutil.copyfile(os.path.join(SRC, f), os.path.join(DST, f))
    cases = json.load(open(os.path.join(SRC, "fit_fixest.json")))
    out = []
    for c in cases:
        if c["name"] not in OLS_CASES:
            continue
        e = {k: c["expect"][k] for k in KEEP if k in c["expect"]}
        out.append({k: c[k] for k in ["name", "data", "y", "x", "fe", "weights", "vcov", "cluster", "ssc"]} | {"expect": e})
    json.dump(out, open(os.path.join(DST, "fixest_ols.json"), "w"), indent=0)
    glm = json.load(open(os.path.join(SRC, "fit_fixest_glm.json")))
    keep = []
    for c in glm:
        if c["name"] not in POIS_CASES:
            continue
        e = {k: c["expect"][k] for k in KEEP + ["deviance", "iterations", "eta", "irls_weights"] if k in c["expect"]}
        keep.append({k: c[k] for k in ["name", "data", "y", "x", "fe", "weights", "offset", "vcov", "cluster", "ssc"]} | {"expect": e})
    json.dump(keep, open(os.path.join(DST, "fixest_pois.json"), "w"), indent=0)
    print(len(out), "ols cases;", len(keep), "poisson cases")


if __name__ == "__main__":
    main()
