"""Subset of the ANSYS .cdb reader of nla3d (FEReaders.cpp:321-729) for the
thermal and solid fixtures: NBLOCK (nodes), EBLOCK (elements with their TYPE
attribute), F (nodal loads), D (nodal fixes), CE (linear constraint equations,
FEReaders.cpp:549-573) and CMBLOCK (node components).

Node numbers are compressed to 1..n in ascending order when they are not already
1..n (MeshData::compressNumbers, FEReaders.cpp:35-86); a number that names no node
maps to 0 there, as the reference's std::map does.
"""
import re

DOF_LABELS = {"UX": 0, "UY": 1, "UZ": 2, "ROTX": 3, "ROTY": 4, "ROTZ": 5, "HDSP": 6, "TEMP": 7, "HEAT": 7}


def _fixed_fields(line, width, count):
    return [line[i * width:(i + 1) * width] for i in range(count)]


def read_cdb(path):
    """Returns dict: nodes {num: (x,y,z)}, cells [(type, [node nums], elnum)], loads [(node, dof, value)],
    fixes [(node, dof, value)], mpcs [([(node, dof, coef), ...], b)], comps {name: [node nums]}, after number
    compression."""
    if str(path).endswith(".gz"):
        import gzip
        with gzip.open(path, "rt") as f:
            lines = f.read().splitlines()
    else:
        with open(path) as f:
            lines = f.read().splitlines()
    nodes, cells, loads, fixes, comps, mpcs = {}, [], [], [], {}, []
    i = 0
    while i < len(lines):
        ln = lines[i]
        tok = [t.strip() for t in ln.split(",")]
        key = tok[0].upper()
        if key == "NBLOCK":
            fmt = lines[i + 1]
            m = re.match(r"\((\d+)i(\d+),(\d+)e(\d+)", fmt.replace(" ", ""))
            ni, wi, nr, wr = (int(v) for v in m.groups())
            i += 2
            while i < len(lines) and not lines[i].strip().startswith("N,R5.3") and lines[i].strip() != "-1":
                s = lines[i]
                ints = [int(v) for v in _fixed_fields(s, wi, ni)]
                rest = s[ni * wi:]
                vals = [float(rest[k * wr:(k + 1) * wr]) for k in range(len(rest) // wr)]
                xyz = (vals + [0.0, 0.0, 0.0])[:3]
                nodes[ints[0]] = tuple(xyz)
                i += 1
            continue
        if key == "EBLOCK":
            fmt = lines[i + 1]
            m = re.match(r"\((\d+)i(\d+)\)", fmt.replace(" ", ""))
            w = int(m.group(2))
            i += 2
            while i < len(lines) and lines[i].strip() != "-1":
                s = lines[i]
                f = [int(v) for v in _fixed_fields(s, w, len(s) // w)]
                etype, nnod, elnum = f[1], f[8], f[10]
                enodes = f[11:11 + nnod]
                cells.append((etype, enodes, elnum))
                i += 1
            continue
        if key == "F" and len(tok) >= 4:
            loads.append((int(tok[1]), DOF_LABELS[tok[2].upper()], float(tok[3])))
        elif key == "D" and len(tok) >= 4:
            fixes.append((int(tok[1]), DOF_LABELS[tok[2].upper()], float(tok[3])))
        elif key == "CE" and len(tok) >= 6 and tok[2].upper() == "DEFI":
            # CE,R5.0,DEFI,nTerms,id,b then CE,R5.0,NODE lines of up to two (node, DOF, coef) terms
            # (the reference reads max(nTerms, 2) terms from each line, which equals what is
            # there only while at most two terms remain; min() is what the format holds)
            left, b, terms = int(tok[3]), float(tok[5]), []
            while left > 0:
                i += 1
                t = [v.strip() for v in lines[i].split(",")]
                for k in range(min(left, 2)):
                    terms.append((int(t[3 + 3 * k]), DOF_LABELS[t[4 + 3 * k].upper()], float(t[5 + 3 * k])))
                left -= min(left, 2)
            mpcs.append((terms, b))
        elif key == "CMBLOCK" and tok[2].upper() == "NODE":
            name, cnt = tok[1], int(tok[3].split("!")[0])
            vals = []
            i += 2
            while len(vals) < cnt:
                vals += [int(v) for v in lines[i].split()]
                i += 1
            # negative entries mark ranges (CMBLOCK "a -b" = a..b)
            out = []
            for k, v in enumerate(vals):
                if v < 0:
                    out += list(range(out[-1] + 1, -v + 1))
                else:
                    out.append(v)
            comps[name] = out
            continue
        i += 1
    # compressNumbers
    if nodes and len(nodes) == max(nodes):
        nmap = None  # numbers already 1..n: nothing to do
    else:
        nmap = {old: k + 1 for k, old in enumerate(sorted(nodes))}

    def m(n):
        return n if nmap is None else nmap.get(n, 0)

    cells.sort(key=lambda c: c[2])
    return {
        "nodes": [nodes[o] for o in sorted(nodes)],
        "cells": [(t, [m(n) for n in ns]) for t, ns, _ in cells],
        "loads": [(m(n), d, v) for n, d, v in loads],
        "fixes": [(m(n), d, v) for n, d, v in fixes],
        "mpcs": [([(m(n), d, c) for n, d, c in terms], b) for terms, b in mpcs],
        "comps": {k: [m(n) for n in v] for k, v in comps.items()},
    }
