"""Free-format MPS/QPS reader (QUADOBJ quadratic section).

Semantics follow what read_mps (src/utils.jl:34-64) receives from QPSReader:
  * the first N row is the objective; other N rows are dropped;
  * a RHS entry on the objective row gives the constant  c0 = -value;
  * default variable bounds are [0, +inf); UP with a negative value and a
    zero lower bound sets the lower bound to -inf; MI/PL/FR/FX/BV/LO/UP;
  * RANGES: E rows become [rhs, rhs+|R|] (R>0) or [rhs-|R|, rhs] (R<0),
    L rows [rhs-|R|, rhs], G rows [rhs, rhs+|R|];
  * QUADOBJ lists the lower triangle once; read_mps then forms
    Q + Q' - Diag(Q) (src/utils.jl:50-53).
Returns (Q, c, A, lcon, ucon, lvar, uvar, c0) with scipy CSC matrices.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def read_mps(path):
    rows = {}          # name -> (index, type)
    row_types = []
    obj_name = None
    cols = {}
    col_names = []
    a_r, a_c, a_v = [], [], []
    c_entries = {}
    rhs = {}
    ranges = {}
    bounds = []
    q_r, q_c, q_v = [], [], []
    c0 = 0.0
    section = None
    with open(path) as f:
        for raw in f:
            line = raw.rstrip("\n")
            if not line.strip() or line.lstrip().startswith("*"):
                continue
            if not line[0].isspace():
                tok = line.split()
                section = tok[0].upper()
                if section in ("QUADOBJ", "QSECTION", "QMATRIX") and section != "QUADOBJ":
                    raise ValueError(f"unsupported MPS section {section}")
                continue
            tok = line.split()
            if section == "ROWS":
                t, name = tok[0].upper(), tok[1]
                if t == "N":
                    if obj_name is None:
                        obj_name = name
                    continue
                rows[name] = (len(row_types), t)
                row_types.append(t)
            elif section == "COLUMNS":
                if "MARKER" in (x.upper().strip("'") for x in tok):
                    continue
                cname = tok[0]
                if cname not in cols:
                    cols[cname] = len(col_names)
                    col_names.append(cname)
                j = cols[cname]
                for k in range(1, len(tok) - 1, 2):
                    rname, val = tok[k], float(tok[k + 1])
                    if rname == obj_name:
                        c_entries[j] = c_entries.get(j, 0.0) + val
                    elif rname in rows:
                        a_r.append(rows[rname][0])
                        a_c.append(j)
                        a_v.append(val)
            elif section == "RHS":
                pairs = tok[1:] if len(tok) % 2 == 1 else tok
                for k in range(0, len(pairs) - 1, 2):
                    rname, val = pairs[k], float(pairs[k + 1])
                    if rname == obj_name:
                        c0 = -val
                    elif rname in rows:
                        rhs[rows[rname][0]] = val
            elif section == "RANGES":
                pairs = tok[1:] if len(tok) % 2 == 1 else tok
                for k in range(0, len(pairs) - 1, 2):
                    rname, val = pairs[k], float(pairs[k + 1])
                    if rname in rows:
                        ranges[rows[rname][0]] = val
            elif section == "BOUNDS":
                bt = tok[0].upper()
                if bt in ("FR", "MI", "PL", "BV"):
                    cname = tok[2] if len(tok) >= 3 else tok[1]
                    bounds.append((bt, cname, None))
                else:
                    bounds.append((bt, tok[2], float(tok[3])))
            elif section == "QUADOBJ":
                i, j, v = cols[tok[0]], cols[tok[1]], float(tok[2])
                q_r.append(i)
                q_c.append(j)
                q_v.append(v)
            elif section == "ENDATA":
                break
    n, m = len(col_names), len(row_types)
    c = np.zeros(n)
    for j, v in c_entries.items():
        c[j] = v
    A = sp.csc_matrix((a_v, (a_r, a_c)), shape=(m, n))
    A.sum_duplicates()
    lcon = np.empty(m)
    ucon = np.empty(m)
    for i, t in enumerate(row_types):
        r = rhs.get(i, 0.0)
        if t == "E":
            lcon[i], ucon[i] = r, r
        elif t == "L":
            lcon[i], ucon[i] = -np.inf, r
        elif t == "G":
            lcon[i], ucon[i] = r, np.inf
        else:
            raise ValueError(f"unknown row type {t}")
        if i in ranges:
            R = ranges[i]
            if t == "E":
                if R > 0:
                    ucon[i] = r + abs(R)
                elif R < 0:
                    lcon[i] = r - abs(R)
            elif t == "L":
                lcon[i] = r - abs(R)
            elif t == "G":
                ucon[i] = r + abs(R)
    lvar = np.zeros(n)
    uvar = np.full(n, np.inf)
    for bt, cname, val in bounds:
        j = cols[cname]
        if bt == "UP":
            uvar[j] = val
            if val < 0 and lvar[j] == 0.0:
                lvar[j] = -np.inf
        elif bt == "LO":
            lvar[j] = val
        elif bt == "FX":
            lvar[j] = uvar[j] = val
        elif bt == "FR":
            lvar[j], uvar[j] = -np.inf, np.inf
        elif bt == "MI":
            lvar[j] = -np.inf
        elif bt == "PL":
            uvar[j] = np.inf
        elif bt == "BV":
            lvar[j], uvar[j] = 0.0, 1.0
        elif bt in ("LI", "UI"):
            (lvar if bt == "LI" else uvar)[j] = val
        else:
            raise ValueError(f"unknown bound type {bt}")
    Q = sp.csc_matrix((q_v, (q_r, q_c)), shape=(n, n))
    Q.sum_duplicates()
    d = Q.diagonal()
    Q = (Q + Q.T - sp.diags(d)).tocsc()
    Q.sum_duplicates()
    return Q, c, A, lcon, ucon, lvar, uvar, c0
