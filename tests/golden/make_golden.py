"""Regenerate the golden fixtures from the reference's in-source unit tests.

The reference (beamform/pacmap-rs 0.2.6) keeps its known-answer vectors inside
``#[cfg(test)]`` modules.  This script reads those Rust files AS TEXT (it runs
nothing from them), pulls out the literal ``array![...]`` values of each test
and writes them as small JSON fixtures next to this file, each tagged with the
file:line it came from.  Only this script reads the reference tree; the tests
load the committed JSON.

    python tests/golden/make_golden.py [/path/to/reference]
"""
from __future__ import annotations

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _src(ref, rel):
    with open(os.path.join(ref, rel), encoding="utf-8") as fh:
        return fh.read()


def _line_of(text, pos):
    return text.count("\n", 0, pos) + 1


def _parse_array(body):
    """Parse a Rust array![...] body (nested brackets, numbers) into lists."""
    body = re.sub(r"//[^\n]*", "", body)
    body = body.replace("\n", " ")
    body = re.sub(r"\b(\d+(?:\.\d*)?(?:[eE][-+]?\d+)?)_?(f32|f64|u32|usize)\b", r"\1", body)
    body = re.sub(r",\s*\]", "]", "[" + body + "]")
    return json.loads(body)


def _find_let_array(text, name, start=0):
    """Find `let [mut] name[: T] = array![ ... ];` from `start`; return (value, line)."""
    m = re.compile(r"let\s+(?:mut\s+)?" + re.escape(name) + r"(?:\s*:\s*[^=]+)?\s*=\s*array!\[").search(
        text, start)
    if not m:
        raise KeyError(name)
    i = m.end()
    depth = 1
    j = i
    while depth:
        if text[j] == "[":
            depth += 1
        elif text[j] == "]":
            depth -= 1
        j += 1
    return _parse_array(text[i:j - 1]), _line_of(text, m.start())


def _find_scalar(text, name, start=0):
    m = re.compile(r"let\s+" + re.escape(name) + r"\s*=\s*([-0-9.eE]+)\s*;").search(text, start)
    if not m:
        raise KeyError(name)
    return float(m.group(1)), _line_of(text, m.start())


def gradient(ref):
    rel = "src/gradient.rs"
    t = _src(ref, rel)
    s = t.index("fn test_pacmap_grad")
    out = {"source": rel}
    for name in ("y_test", "pair_neighbors", "pair_mn", "pair_fp", "grad_python"):
        val, line = _find_let_array(t, name, s)
        out[name] = val
        out[name + "_line"] = line
    for name in ("w_neighbors", "w_mn", "w_fp"):
        val, line = _find_scalar(t, name, s)
        out[name] = val
    out["epsilon"] = 1e-6
    return out


def adam(ref):
    rel = "src/adam.rs"
    t = _src(ref, rel)
    s = t.index("fn test_update_embedding_adam")
    out = {"source": rel}
    for name in ("y", "grad", "y_expected", "m_expected", "v_expected"):
        val, line = _find_let_array(t, name, s)
        out[name] = val
        out[name + "_line"] = line
    for name in ("beta1", "beta2", "lr", "itr"):
        val, _ = _find_scalar(t, name, s)
        out[name] = val
    out["epsilon"] = 1e-6
    return out


def weights(ref):
    rel = "src/weights.rs"
    t = _src(ref, rel)
    s = t.index("fn test_find_weight")
    w_init, _ = _find_scalar(t, "w_mn_init", s)
    cases = []
    pat = re.compile(r"let\s+(w\d+)\s*=\s*find_weights\(w_mn_init,\s*(\d+),\s*(\d+),\s*(\d+)\);")
    for m in pat.finditer(t, s):
        var = m.group(1)
        exp = {}
        for field in ("w_mn", "w_neighbors", "w_fp"):
            mm = re.search(r"assert_abs_diff_eq!\(" + var + r"\." + field + r",\s*([-0-9.]+)\)", t[m.end():])
            exp[field] = float(mm.group(1))
        cases.append({"itr": int(m.group(2)), "p1": int(m.group(3)), "p2": int(m.group(4)),
                      "expected": exp, "line": _line_of(t, m.start())})
    return {"source": rel, "w_mn_init": w_init, "cases": cases}


def scale_dist(ref):
    rel = "src/distance.rs"
    t = _src(ref, rel)
    out = {"source": rel, "cases": []}
    for fn, tol in (("fn test_scale_dist()", 0.0), ("fn test_scale_dist_large()", 1e-6)):
        s = t.index(fn)
        case = {"name": fn.split()[1].rstrip("()"), "tolerance": tol}
        for name in ("knn_distances", "sig", "neighbors", "expected"):
            val, line = _find_let_array(t, name, s)
            case[name] = val
            case[name + "_line"] = line
        out["cases"].append(case)
    return out


def knn(ref):
    rel = "src/knn.rs"
    t = _src(ref, rel)
    out = {"source": rel, "cases": []}
    sqrt2 = 2 ** 0.5
    for fn in ("fn test_normal_case", "fn test_duplicate_embeddings", "fn test_negative_components",
               "fn test_k_exceeds", "fn test_k_zero", "fn test_single_embedding"):
        s = t.index(fn)
        emb, line = _find_let_array(t, "embeddings", s)
        k = int(re.search(r"let\s+k\s*=\s*(\d+)\s*;", t[s:]).group(1))
        shp = re.search(r"assert_eq!\(neighbor_indices\.shape\(\),\s*&\[(\d+),\s*(\d+)\]\)", t[s:])
        case = {"name": fn.split()[1], "embeddings": emb, "k": k, "line": line,
                "shape": [int(shp.group(1)), int(shp.group(2))]}
        em = re.search(r"let\s+expected\s*=\s*vec!\[(.*?)\];", t[s:s + 3000], re.S)
        if em and t[s:].index("let expected") < t[s + 10:].find("fn ") + 10:
            rows = []
            for rm in re.finditer(r"\((\d+),\s*vec!\[(.*?)\]\)", em.group(1)):
                nb = []
                for pm in re.finditer(r"\((\d+),\s*([A-Za-z0-9_.]+)\)", rm.group(2)):
                    v = pm.group(2)
                    val = 1 / sqrt2 if v == "FRAC_1_SQRT_2" else float(v)
                    nb.append([int(pm.group(1)), val])
                rows.append([int(rm.group(1)), nb])
            case["expected"] = rows
            case["tolerance"] = {"epsilon": 1e-4, "max_relative": 1e-4}
        out["cases"].append(case)
    # knn.rs test_empty_embeddings: zeros((0, 128)), k=5 -> shapes [0, 0]
    s = t.index("fn test_empty_embeddings")
    out["cases"].append({"name": "test_empty_embeddings", "embeddings_shape": [0, 128], "k": 5,
                         "shape": [0, 0], "line": _line_of(t, s)})
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    fixtures = {"gradient.json": gradient(ref), "adam.json": adam(ref), "weights.json": weights(ref),
                "scale_dist.json": scale_dist(ref), "knn.json": knn(ref)}
    for name, data in fixtures.items():
        with open(os.path.join(HERE, name), "w", encoding="utf-8") as fh:
            json.dump(data, fh, indent=1)
            fh.write("\n")
        print("wrote", name)


if __name__ == "__main__":
    main()
