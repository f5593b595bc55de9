"""Extracts golden vectors from the reference's own test fixtures into tests/golden/.

Run in the build container only (it reads the read-only reference checkout); the outputs are
small committed data files, so nothing at run time reads the reference.

Sources (fixtures of crates/fereg-core/tests, written there by fixest 0.14.2):
  tests/data/fit_d1.csv, fit_toy.csv          the input panels
  tests/data/fit_fixest.json                  OLS cases (tests/fit_ols.rs:10-67, RTOL 1e-8)
  tests/data/fit_fixest_glm.json              Poisson cases (tests/fit_glm.rs)
  tests/data/fit_se.csv, fit_se_fixest.json   HC2 / HC3 / CRV3 cases (fixest 0.14.2 through
                                              tests/data/gen_se_fixtures.R; OLS and Poisson,
                                              intercept-only absorbed terms, no IV)
  tests/data/fit_fixest_fe.json (sl_*),       varying slopes (fe[x] terms, with and without the
  fit_fixest_glm.json (pois_sl*),             level intercepts): OLS and Poisson on fit_d1.csv,
  fit_se_fixest.json (*_sl)                   HC2 / HC3 / CRV3 / hetero on fit_se.csv
  tests/data/ols_data.csv, ols_s*.json        reghdfe 6.13.1 ground truth (tests/ols.rs:765-935,
                                              written by tests/data/gen_ols_fixtures.do); the
                                              model of each fixture restates the table of
                                              tests/ols.rs:803-838
"""
import json
import os
import shutil

SRC = "/root/reference/crates/fereg-core/tests/data"
DST = os.path.dirname(os.path.abspath(__file__))

OLS_CASES = None  # every OLS case
POIS_CASES = ["pois_fe2_iid", "pois_fe2_hetero", "pois_collin", "pois_inter_fx", "pois_fe2_off_w", "pois_fe2_cl_c3_h"]
KEEP = ["coef", "se", "vcov", "df_k", "nobs", "nparams", "collin", "obs", "G", "resid", "ssr", "coef_names"]

# reghdfe fixtures: (name, y, x, absorb terms, cluster dims, robust, weight (var, kind), singletons,
# noconstant, dof options (firstpair, pairwise, clusters, continuous), expected status)
_X3, _X2 = ["x1", "x2", "x3"], ["x1", "x2"]
_IDYEAR = [["id"], ["year"]]
_DOF = [False, True, True, True]
REGHDFE = [
    ("s01_unadj", "y", _X3, _IDYEAR, [], False, None, True, False, _DOF, []),
    ("s02_robust_3fe", "y", _X3, [["id"], ["year"], ["firm"]], [], True, None, True, False, _DOF, []),
    ("s03_cluster_eq_fe", "y", _X2, _IDYEAR, [["id"]], False, None, True, False, _DOF, []),
    ("s04_fe_nested_in_cluster", "y", _X3, _IDYEAR, [["region"]], False, None, True, False, _DOF, []),
    ("s05_twoway_psd", "y", _X3, [["firm"], ["year"]], [["c1"], ["c2"]], False, None, True, False, _DOF, []),
    ("s06_aw_cluster", "y", _X3, _IDYEAR, [["firm"]], False, ["w", "aweight"], True, False, _DOF, []),
    ("s07_fw_robust", "y", _X3, _IDYEAR, [], True, ["fw", "fweight"], True, False, _DOF, []),
    ("s08_pw", "y", _X3, [["id"], ["year"], ["firm"]], [], False, ["w", "pweight"], True, False, _DOF, []),
    ("s09_absorbed_collinear", "y", ["x1", "x2", "xa", "xc"], _IDYEAR, [], False, None, True, False, _DOF, [0, 0, 2, 3]),
    ("s10_interaction", "y", _X2, [["id"], ["firm", "year"]], [], False, None, True, False, _DOF, []),
    ("s11_noabsorb_cluster", "y", _X3, [], [["firm"]], False, None, True, False, _DOF, []),
    ("s12_noconstant", "y", _X2, _IDYEAR, [], False, None, True, True, _DOF, []),
    ("s13_keepsing_twoway", "y", _X2, _IDYEAR, [["id"], ["year"]], False, None, False, False, _DOF, []),
    ("s14_groupvar", "y", _X2, [["id"], ["firm"], ["year"]], [], False, None, True, False, _DOF, []),
    ("s15_firstpair_cluster", "y", _X2, [["id"], ["year"], ["firm"]], [["firm"]], False, None, True, False,
     [True, False, False, False], []),
    ("s17_y_absorbed", "y2", _X2, _IDYEAR, [], False, None, True, False, _DOF, []),
    ("s18_fw_twoway", "y", _X3, [["firm"], ["year"]], [["c1"], ["c2"]], False, ["fw", "fweight"], True, False, _DOF, []),
]
REGHDFE_KEEP = ["cmdline", "colnames", "b", "V", "dof_table", "sample", "scalars", "gv14"]


def main():
    for f in ["fit_d1.csv", "fit_toy.csv"]:
        shutil.copyfile(os.path.join(SRC, f), os.path.join(DST, f))
    cases = json.load(open(os.path.join(SRC, "fit_fixest.json")))
    out = []
    for c in cases:
        if OLS_CASES is not None and c["name"] not in OLS_CASES:
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
    shutil.copyfile(os.path.join(SRC, "ols_data.csv"), os.path.join(DST, "ols_data.csv"))
    rg = []
    for name, y, x, absorb, vce, robust, weight, sing, nocons, dof, status in REGHDFE:
        j = json.load(open(os.path.join(SRC, f"ols_{name}.json")))
        rg.append({"name": name, "y": y, "x": x, "absorb": absorb, "cluster": vce, "robust": robust,
                   "weight": weight, "singletons": sing, "noconstant": nocons, "dof": dof, "status": status,
                   "expect": {k: j[k] for k in REGHDFE_KEEP if k in j}})
    json.dump(rg, open(os.path.join(DST, "reghdfe_ols.json"), "w"))
    shutil.copyfile(os.path.join(SRC, "fit_se.csv"), os.path.join(DST, "fit_se.csv"))
    nrows = sum(1 for _ in open(os.path.join(SRC, "fit_se.csv"))) - 1
    se = []
    for c in json.load(open(os.path.join(SRC, "fit_se_fixest.json"))):
        if c["vcov"] not in ("hc2", "hc3", "crv3") or c["model"] not in ("ols", "poisson") or c["endog"]:
            continue
        if any(t["slopes"] or not t["intercept"] for t in c["fe"]) or c["name"].startswith("iv_"):
            continue
        e = {k: c["expect"][k] for k in ["coef", "se", "vcov", "df_k", "nobs", "collin", "G", "coef_names"]
             if k in c["expect"]}
        if e["nobs"] != nrows:
            continue  # every case here keeps all rows; the fixture carries no row list
        e["obs"] = list(range(1, nrows + 1))
        se.append({k: c[k] for k in ["name", "data", "model", "y", "x", "fe", "intercept", "weights", "vcov",
                                      "cluster", "ssc"]} | {"expect": e})
    json.dump(se, open(os.path.join(DST, "fixest_se.json"), "w"), indent=0)
    sl = []
    for src, model in (("fit_fixest_fe.json", "ols"), ("fit_fixest_glm.json", "poisson")):
        for c in json.load(open(os.path.join(SRC, src))):
            if not any(t["slopes"] or not t["intercept"] for t in c["fe"]) or c.get("endog"):
                continue
            e = {k: c["expect"][k] for k in KEEP + ["deviance", "iterations", "fitted"] if k in c["expect"]}
            sl.append({k: c[k] for k in ["name", "data", "y", "x", "fe", "weights", "vcov", "cluster", "ssc"]}
                      | {"model": model, "offset": c.get("offset"), "expect": e})
    nse = sum(1 for _ in open(os.path.join(SRC, "fit_se.csv"))) - 1
    for c in json.load(open(os.path.join(SRC, "fit_se_fixest.json"))):
        if not any(t["slopes"] or not t["intercept"] for t in c["fe"]) or c.get("endog"):
            continue
        if c["vcov"] not in ("hc2", "hc3", "crv3", "hetero") or c["model"] not in ("ols", "poisson"):
            continue
        e = {k: c["expect"][k] for k in ["coef", "se", "vcov", "df_k", "nobs", "collin", "G", "coef_names"]
             if k in c["expect"]}
        if e["nobs"] != nse:
            continue  # every case here keeps all rows; the fixture carries no row list
        e["obs"] = list(range(1, nse + 1))
        sl.append({k: c[k] for k in ["name", "data", "y", "x", "fe", "weights", "vcov", "cluster", "ssc"]}
                  | {"model": c["model"], "offset": None, "expect": e})
    json.dump(sl, open(os.path.join(DST, "fixest_slopes.json"), "w"), indent=0)
    print(len(out), "ols cases;", len(sl), "slope cases;", len(keep), "poisson cases;", len(rg), "reghdfe cases")


if __name__ == "__main__":
    main()
