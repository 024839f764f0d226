"""Subset of the ANSYS .cdb reader of nla3d (FEReaders.cpp:321-729) for the
thermal and solid fixtures: NBLOCK (nodes), EBLOCK (elements with their TYPE
attribute), F (nodal loads), D (nodal fixes) and CMBLOCK (node components).

Numbers are compressed to 1..n in ascending order (MeshData::compressNumbers,
FEReaders.cpp:35-86).
"""
import re

DOF_LABELS = {"UX": 0, "UY": 1, "UZ": 2, "ROTX": 3, "ROTY": 4, "ROTZ": 5, "HDSP": 6, "TEMP": 7, "HEAT": 7}


def _fixed_fields(line, width, count):
    return [line[i * width:(i + 1) * width] for i in range(count)]


def read_cdb(path):
    """Returns dict: nodes {num: (x,y,z)}, cells [(type, [node nums], elnum)], loads [(node, dof, value)],
    fixes [(node, dof, value)], comps {name: [node nums]}, after number compression."""
    with open(path) as f:
        lines = f.read().splitlines()
    nodes, cells, loads, fixes, comps = {}, [], [], [], {}
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
    nmap = {old: k + 1 for k, old in enumerate(sorted(nodes))}
    cells.sort(key=lambda c: c[2])
    return {
        "nodes": [nodes[o] for o in sorted(nodes)],
        "cells": [(t, [nmap[n] for n in ns]) for t, ns, _ in cells],
        "loads": [(nmap[n], d, v) for n, d, v in loads],
        "fixes": [(nmap[n], d, v) for n, d, v in fixes],
        "comps": {k: [nmap[n] for n in v] for k, v in comps.items()},
    }
