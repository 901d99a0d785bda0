"""Writes flop_constants.json: the frozen algorithmic FLOP counts per unit of work that the
benchmark's roofline numerator uses, so that the figure is comparable from one build to the next.

Two formulations are counted once and frozen:

* reference: the reference's own formulation as written.  FSYS, JVAR and JPAR
  (htpykinematics.cu:8-1384) are counted from their source text: every cplxAdd/cplxSub counts 2
  real operations, cplxMul 6, cplxDiv 9 (Smith's division), cplxPow(x, k) (k - 1) multiplications
  (support.cu:5-15), and cplxMul({ -1, 0 }, .) 0 (a sign flip); the per-pose loop body counts
  seven times.  q and dqdt (htpysupport.cu:37-62), the dense JPAR . dq/dt product (Matr::mul,
  classes.cuh:190-199, htpysupport.cu:121), the row scaling (htpysupport.cu:260-277, :446-471),
  LinearSolve and LinearSolveNewton (htpysupport.cu:146-232, :302-400), the tangent
  normalisation (:285-296), the Newton update and residual (:475-491) and the RK6(5)8M stage
  combinations, error estimate and solution (rungekutta.cu:43-122) are counted from their loop
  bounds, with the same conversion.
* compact: this implementation's formulation, as counted by frg_flops() (the per-pose product
  rule, the directional dF/dt and the structured solve).

A path's flops are tangent_solves * tangent_solve + newton_its * newton_update
+ rk_attempts * rk_combination, from the path's own counters.

    python finite_root_generation_amd/flop_constants.py      # needs the reference sources
"""
from __future__ import annotations

import json
import os
import re
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REF = os.environ.get("FRG_REFERENCE", "/root/reference")
OUT = os.path.join(PKG, "flop_constants.json")
NV, NP, NY = 21, 30, 23
CADD, CMUL, CDIV, CFMA = 2, 6, 9, 8


def _function_body(text: str, name: str) -> str:
    m = re.search(r"__device__\s+void\s+" + name + r"\s*\(", text)
    i = text.index("{", m.end())
    depth, j = 0, i
    while True:
        if text[j] == "{":
            depth += 1
        elif text[j] == "}":
            depth -= 1
            if depth == 0:
                return text[i:j + 1]
        j += 1


def _strip_comments(s: str) -> str:
    s = re.sub(r"/\*.*?\*/", "", s, flags=re.S)
    s = re.sub(r"//[^\n]*", "", s)
    return s.replace("\\\n", " ")  # the generated expressions continue lines with a backslash


def _census(s: str) -> dict:
    neg = len(re.findall(r"cplxMul\(\s*\{\s*-1\s*,\s*0\s*\}", s))
    mul = len(re.findall(r"cplxMul\(", s)) - neg
    add = len(re.findall(r"cplxAdd\(", s))
    sub = len(re.findall(r"cplxSub\(", s))
    div = len(re.findall(r"cplxDiv\(", s))
    pow_mults = 0
    for m in re.finditer(r"cplxPow\(", s):
        # the exponent is the last argument: "{ k, 0 })" or "k)"
        depth, j = 1, m.end()
        while depth:
            depth += {"(": 1, ")": -1}.get(s[j], 0)
            j += 1
        args = s[m.end():j - 1]
        k = re.search(r",\s*\{?\s*(\d+)\s*(?:,\s*0\s*\})?\s*$", args)
        pow_mults += int(k.group(1)) - 1
    flops = CADD * (add + sub) + CMUL * (mul + pow_mults) + CDIV * div
    return {"add": add, "sub": sub, "mul": mul, "neg": neg, "div": div, "pow_mults": pow_mults, "flops": flops}


def system_census(path: str) -> dict:
    text = _strip_comments(open(path).read())
    out = {}
    for name in ("FSYS", "JVAR", "JPAR"):
        body = _function_body(text, name)
        loop = body.index("for (int j = 1; j <= 7; ++j)")
        pre, inner = _census(body[:loop]), _census(body[loop:])
        out[name] = {"per_pose": inner, "flops": pre["flops"] + 7 * inner["flops"]}
    return out


def lu_flops(rows: int, cols: int, pivot_rows: int) -> int:
    """LinearSolve / LinearSolveNewton: pivot search |a|^2 (3), per eliminated row one division
    and (cols - i - 1) fused updates (mul + sub) plus the right-hand side; back substitution with
    cplxFma, 1/a_ii and one product per unknown solved."""
    f = 0
    for i in range(cols):
        f += 3 * max(0, pivot_rows - i)
        for _ in range(i + 1, pivot_rows):
            f += CDIV + (cols - i - 1) * (CMUL + CADD) + (CMUL + CADD)
    for i in range(rows):
        f += (cols - 1 - i) * CFMA + CDIV + CADD + CMUL
    return f


def reference_formulation(ref_dir: str) -> dict:
    sysc = system_census(os.path.join(ref_dir, "htpykinematics.cu"))
    f_q = NP * (3 * CMUL + CADD + CMUL + CADD + CDIV) + 2 * NP           # q: 30 divisions, 1 - t
    f_dqdt = NP * (CADD + CMUL + CMUL + CADD + CMUL + CDIV) + NP         # dqdt: pow 2, division
    f_matvec = NV * NP * (CMUL + CADD)                                   # JPAR . dq/dt
    f_jh = f_q + sysc["JVAR"]["flops"] + sysc["JPAR"]["flops"] + f_dqdt + f_matvec
    f_scale = (NV + 1) * (1 + 2 * NY)                                    # 1/max, 23 scalings per row
    f_norm = NY * 4 + 1
    f_tangent = f_jh + f_scale + lu_flops(NY, NY, NY) + (3 + f_norm + 1 + CMUL + NY * CMUL)
    f_h = f_q + sysc["FSYS"]["flops"] + NY * (CMUL + CADD) + CADD
    f_newton = f_jh + f_h + f_scale + (NV + 1) * 2 + lu_flops(NY - 1, NY, NY - 1) + NY * CADD + f_norm
    terms = [1, 2, 3, 4, 5, 6, 6]                                        # stages 2..8
    vscale, vadd = 1 + NY * CMUL, NY * CADD
    f_rk = sum(t * (vscale + vadd) for t in terms)
    f_rk += 7 * vscale + 6 * vadd + f_norm + 3 + 7 * (vscale + vadd)    # error, norm, factor, solution
    return {"tangent_solve": f_tangent, "newton_update": f_newton, "rk_combination": f_rk,
            "parts": {"q": f_q, "dqdt": f_dqdt, "jpar_dqdt_product": f_matvec, "jh": f_jh, "h": f_h,
                      "row_scaling": f_scale, "linear_solve": lu_flops(NY, NY, NY),
                      "linear_solve_newton": lu_flops(NY - 1, NY, NY - 1)},
            "census": sysc}


def compact_formulation() -> dict:
    sys.path.insert(0, os.path.dirname(PKG))
    from finite_root_generation_amd import capi
    m = capi.flop_model()
    return {k: m[k] for k in ("tangent_solve", "newton_update", "rk_combination", "pose_rows", "pose_values")}


def main() -> None:
    data = {
        "unit": "real floating-point operations per counted event of one path",
        "path_flops": "tangent_solves * tangent_solve + newton_its * newton_update + rk_attempts * rk_combination",
        "reference": reference_formulation(REF),
        "compact": compact_formulation(),
        "conversion": {"complex_add": CADD, "complex_mul": CMUL, "complex_div": CDIV, "complex_fma": CFMA,
                       "sign_flip": 0},
    }
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print(json.dumps({k: {kk: data[k][kk] for kk in ("tangent_solve", "newton_update", "rk_combination")}
                      for k in ("reference", "compact")}))


if __name__ == "__main__":
    main()
