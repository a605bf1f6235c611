"""Extract golden vectors from the reference's own refout files into tests/golden/.

Run in a container where the reference tree is mounted:
    python tests/golden/make_fixtures.py /root/reference
The outputs are data (numbers printed by the reference's tests); the tests read
only these fixtures, never the reference tree.
"""
from __future__ import annotations

import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _floats(line):
    return [float(t) for t in re.findall(r"[-+]?\d+\.\d+(?:e[-+]?\d+)?", line)]


def jacobi3_explicit(path, keep_rows):
    """Interpolation / gradient / mass tables of jacobi-ex3 (src/jacobi/tests/ex3.c:58-107)."""
    lines = open(path).read().splitlines()
    out, i, elem = {}, 0, 0
    while i < len(lines):
        m = re.match(r"Interpolation matrix of size \((\d+),(\d+)\)", lines[i])
        if not m:
            i += 1
            continue
        Q, P = int(m.group(1)), int(m.group(2))
        interp = np.array([_floats(l) for l in lines[i + 1:i + 1 + Q]])
        i += 1 + Q
        assert lines[i].startswith("Gradient matrix")
        grad = np.array([_floats(l) for l in lines[i + 1:i + 1 + Q]])
        i += 1 + Q
        assert lines[i].startswith("Mass matrix")
        mass = np.array([_floats(l) for l in lines[i + 1:i + 1 + P]])
        i += 1 + P
        r = min(Q, keep_rows[elem])
        out[f"e{elem}_interp"] = interp[:r].astype(np.float32)
        out[f"e{elem}_grad"] = grad[:r].astype(np.float32)
        out[f"e{elem}_mass"] = mass.astype(np.float32)
        out[f"e{elem}_shape"] = np.array([Q, P, r])
        elem += 1
    return out


def jacobi3_sparse(path):
    """Sparse Q1 tables of jacobi-ex3 -sparse (src/jacobi/tests/ex3.c:109-140)."""
    data = {}
    for line in open(path):
        m = re.match(r"elem (\d+)(?:  patch (\d+))?\s*:?\s*(\w+)(?:\(patch,\w\))?\[\s*(\d+)\]\[\s*\d+:\s*\d+\]\s+(.*)", line)
        if not m:
            continue
        e, patch, kind, row, rest = m.groups()
        vals = [float(t) for t in rest.split()]
        key = (int(e), kind, int(patch) if patch is not None else -1)
        data.setdefault(key, []).append((int(row), vals))
    out = {}
    for (e, kind, patch), rows in data.items():
        rows.sort()
        arr = np.array([v for _, v in rows])
        name = f"e{e}_{kind}" + (f"_p{patch}" if patch >= 0 else "")
        out[name] = arr.astype(np.int32) if kind in ("qidx", "bidx") else arr.astype(np.float32)
    return out


def jacobi4_patches(path):
    """Patch indices and reported weights of jacobi-ex4 (src/jacobi/tests/ex4.c:7-46)."""
    out, elem = {}, -1
    for line in open(path):
        m = re.match(r"Element (\d+): npatches (\d+)\s+patchsize (\d+)", line)
        if m:
            elem = int(m.group(1))
            out[f"e{elem}_np"] = np.array([int(m.group(2)), int(m.group(3))])
            continue
        m = re.match(r"\s*(indices|weights)\[\s*(\d+)\]\[.*?\]\s+(.*)", line)
        if m:
            kind, row, rest = m.groups()
            key = f"e{elem}_{kind}"
            vals = [float(t) for t in rest.split()]
            out.setdefault(key, []).append(vals)
    return {k: (np.array(v, dtype=np.int32) if k.endswith("indices") else np.array(v, dtype=np.float64))
            for k, v in out.items()}


def ellip_history(path):
    """SNES function norms and KSP residual histories of an ellip run (src/fs/tests/ellip.c)."""
    snes, ksp, cur = [], [], None
    norms = {}
    for line in open(path):
        m = re.match(r"Pointwise (solution|gradient) error\s+\|x\|_1 (\S+)\s+\|x\|_2 (\S+)\s+\|x\|_inf (\S+)", line)
        if m:
            norms[m.group(1)] = [float(m.group(k)) for k in (2, 3, 4)]
            continue
        m = re.match(r"\s*(\d+) SNES Function norm (<)?\s*([\d.e+-]+)", line)
        if m:
            snes.append(float(m.group(3)) if not m.group(2) else 0.0)
            cur = []
            ksp.append(cur)
            continue
        m = re.match(r"\s*(\d+) KSP Residual norm (<)?\s*([\d.e+-]+)", line)
        if m:
            if cur is None:
                cur = []
                ksp.append(cur)
            cur.append(float(m.group(3)) if not m.group(2) else 0.0)
    out = {"snes": snes, "ksp": [k for k in ksp if k]}
    if norms:
        out["error_norms"] = norms
    return out


def main(ref):
    jt = os.path.join(ref, "src/jacobi/tests/refout")
    ft = os.path.join(ref, "src/fs/tests/refout")
    np.savez_compressed(os.path.join(HERE, "jacobi3_tensor_gauss.npz"),
                        **jacobi3_explicit(os.path.join(jt, "jacobi-3-tensor-gauss.refout"), [125, 125, 24, 12]))
    np.savez_compressed(os.path.join(HERE, "jacobi3_tensor_lobatto.npz"),
                        **jacobi3_explicit(os.path.join(jt, "jacobi-3-tensor-lobatto.refout"), [125, 125, 24, 12]))
    for r in (4, 6):
        np.savez_compressed(os.path.join(HERE, f"jacobi3_sparsex_r{r}.npz"),
                            **jacobi3_sparse(os.path.join(jt, f"jacobi-3-tensor-sparsex-r{r}.refout")))
    for m in ("fast", "sparse"):
        np.savez_compressed(os.path.join(HERE, f"jacobi4_{m}.npz"),
                            **jacobi4_patches(os.path.join(jt, f"jacobi-4-tensor-{m}.refout")))
    hist = {}
    for name in ("ellip-e0-b1-p16", "ellip-e0-b3-p16", "ellip-e0-b3-p14-sse", "ellip-e0-b3-p17-lam1-sse",
                 "ellip-morph", "ellip-morph-qfast"):
        hist[name] = ellip_history(os.path.join(ft, name + ".refout"))
    with open(os.path.join(HERE, "ellip_histories.json"), "w") as f:
        json.dump(hist, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
