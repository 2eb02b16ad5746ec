"""Reader and writer for the Abaqus-style `.inp` subset the linear path needs.

Semantics follow crates/solvix-io/src/parser.rs for the keywords used by the
path's decks: *NODE (146-173: deck order, z = 0 for 2-field records),
*ELEMENT TYPE=CPS3|CPS4|C3D8 [ELSET=] (175-256, trailing-comma continuation),
*NSET / *ELSET [GENERATE] (258-312, names upper-cased, duplicates kept),
*MATERIAL / *ELASTIC (496-535), *SOLID SECTION (537-584, model-global
thickness), *STEP (1114-1150: a step starts from the previous step's state),
*BOUNDARY (742-796: node/set, dof[, dof2]), *CLOAD (798-844, expanded over
sets) and the single-step mirroring of 1325-1336. Other keywords are skipped.
"""
from __future__ import annotations

from typing import Dict, List

import numpy as np

from . import _abi
from .model import Model


class ParseError(Exception):
    def __init__(self, line: int, message: str):
        super().__init__(f"line {line}: {message}")
        self.line = line
        self.message = message


def _fields(text: str) -> List[str]:
    f = [s.strip() for s in text.split(",")]
    while f and f[-1] == "":
        f.pop()
    return f


def _kw(text: str):
    parts = [p.strip() for p in text[1:].split(",")]
    name = parts[0].lower()
    params = {}
    for p in parts[1:]:
        if not p:
            continue
        if "=" in p:
            k, v = p.split("=", 1)
            params[k.strip().lower()] = v.strip()
        else:
            params[p.strip().lower()] = ""
    return name, params


def parse_inp(text: str, name: str = "model") -> Model:
    lines = text.splitlines()
    items = []  # (kind, lineno, text)
    for i, raw in enumerate(lines, 1):
        s = raw.strip()
        if not s or s.startswith("**"):
            continue
        items.append(("kw" if s.startswith("*") else "data", i, s))

    node_ids: List[int] = []
    coords: List[List[float]] = []
    elem_ids: List[int] = []
    kinds: List[int] = []
    conn: List[List[int]] = []
    nsets: Dict[str, List[int]] = {}
    elsets: List[list] = []  # [name, list]
    materials: List[list] = []  # [name, E, nu]
    assignments: List[tuple] = []
    thickness = 1.0
    pending_material = None
    model_state = {"bcs": [], "cloads": []}
    steps: List[dict] = []
    cur = None

    def upsert_elset(nm, ids):
        key = nm.upper()
        for s in elsets:
            if s[0] == key:
                s[1].extend(ids)
                return
        elsets.append([key, list(ids)])

    def resolve(target, line):
        key = target.upper()
        if key in nsets:
            return list(nsets[key])
        try:
            return [int(target)]
        except ValueError:
            raise ParseError(line, f"unknown node set or node id '{target}'")

    j = 0
    while j < len(items):
        kind, line, text = items[j]
        if kind == "data":
            j += 1
            continue
        kw, params = _kw(text)
        j += 1
        data = []
        while j < len(items) and items[j][0] == "data":
            data.append(items[j][1:])
            j += 1
        if kw == "node":
            for dno, t in data:
                f = _fields(t)
                if len(f) not in (3, 4):
                    raise ParseError(dno, f"*NODE expects id, x, y[, z]; got {len(f)} field(s)")
                node_ids.append(int(f[0]))
                coords.append([float(f[1]), float(f[2]), float(f[3]) if len(f) == 4 else 0.0])
        elif kw == "element":
            ty = params.get("type", "").upper()
            if ty not in _abi.KIND_NAMES:
                continue
            k = _abi.KIND_NAMES[ty]
            want = 1 + _abi.NODES_PER_KIND[k]
            q = 0
            while q < len(data):
                f = []
                first = data[q][0]
                while q < len(data):
                    t = data[q][1]
                    ends_comma = t.endswith(",")
                    f.extend(s.strip() for s in t.split(","))
                    q += 1
                    if not ends_comma or sum(1 for s in f if s) >= want:
                        break
                f = [s for s in f if s]
                if len(f) != want:
                    raise ParseError(first, f"*ELEMENT expects {want} field(s); got {len(f)}")
                eid = int(f[0])
                elem_ids.append(eid)
                kinds.append(k)
                conn.append([int(s) for s in f[1:]])
                if "elset" in params:
                    upsert_elset(params["elset"], [eid])
        elif kw in ("elset", "nset"):
            nm = params.get("elset", params.get("nset"))
            if nm is None:
                continue
            gen = "generate" in params
            for dno, t in data:
                f = _fields(t)
                if gen:
                    start, end = int(f[0]), int(f[1])
                    step = int(f[2]) if len(f) == 3 else 1
                    ids = list(range(start, end + 1, step))
                else:
                    ids = [int(s) for s in f]
                if kw == "elset":
                    upsert_elset(nm, ids)
                else:
                    nsets.setdefault(nm.upper(), []).extend(ids)
        elif kw == "material":
            if "name" in params:
                pending_material = params["name"].upper()
        elif kw == "elastic":
            if data and pending_material is not None:
                f = _fields(data[0][1])
                materials.append([pending_material, float(f[0]), float(f[1])])
                pending_material = None
        elif kw == "solid section":
            es, mat = params.get("elset"), params.get("material")
            if es and mat:
                assignments.append((es.upper(), mat.upper()))
                if data:
                    thickness = float(_fields(data[0][1])[0])
        elif kw == "step":
            if cur is not None:
                steps.append(cur)
            op_new = params.get("op", "").lower() == "new"
            base = {"bcs": [], "cloads": []} if op_new else (steps[-1] if steps else model_state)
            cur = {"bcs": list(base["bcs"]), "cloads": list(base["cloads"])}
        elif kw == "end step":
            if cur is not None:
                steps.append(cur)
                cur = None
        elif kw == "boundary":
            target = cur if cur is not None else model_state
            if cur is not None and params.get("op", "").lower() == "new":
                target["bcs"] = []
            for dno, t in data:
                f = _fields(t)
                nodes = resolve(f[0], dno)
                d1 = int(f[1])
                dofs = list(range(d1, int(f[2]) + 1)) if len(f) >= 3 else [d1]
                target["bcs"].append((nodes, dofs))
        elif kw == "cload":
            target = cur if cur is not None else model_state
            if cur is not None and params.get("op", "").lower() == "new":
                target["cloads"] = []
            for dno, t in data:
                f = _fields(t)
                if len(f) != 3:
                    raise ParseError(dno, f"*CLOAD expects node/set, dof, value; got {len(f)} field(s)")
                for nid in resolve(f[0], dno):
                    target["cloads"].append((nid, int(f[1]), float(f[2])))
    if cur is not None:
        steps.append(cur)
    state = steps[-1] if steps else model_state

    conn_ptr = np.zeros(len(conn) + 1, np.int64)
    conn_ptr[1:] = np.cumsum([len(c) for c in conn])
    cl = state["cloads"]
    return Model(
        node_ids=np.asarray(node_ids, np.uint32),
        coords=np.asarray(coords, np.float64).reshape(-1, 3),
        elem_ids=np.asarray(elem_ids, np.uint32),
        kinds=np.asarray(kinds, np.uint8),
        conn_ptr=conn_ptr,
        conn_ids=np.asarray([n for c in conn for n in c], np.uint32),
        materials=[(m[0], m[1], m[2]) for m in materials],
        elsets=[(s[0], np.asarray(s[1], np.uint32)) for s in elsets],
        assignments=assignments,
        thickness=thickness,
        bcs=[(np.asarray(n, np.uint32), np.asarray(d, np.int32)) for n, d in state["bcs"]],
        cload_node=np.asarray([c[0] for c in cl], np.uint32),
        cload_dof=np.asarray([c[1] for c in cl], np.int32),
        cload_value=np.asarray([c[2] for c in cl], np.float64),
        name=name,
    )


def read_inp(path: str) -> Model:
    import os
    with open(path) as fh:
        return parse_inp(fh.read(), os.path.splitext(os.path.basename(path))[0])


_TYPE_OF_KIND = {v: k for k, v in _abi.KIND_NAMES.items()}


def write_inp(model: Model) -> str:
    """Deck text for a single-material model (layout of crates/solvix-tools/src/deck.rs:12-161)."""
    out = [f"*HEADING\n{model.name}", "*NODE"]
    ndim = 3 if model.ndof() == 3 else 2
    for i, nid in enumerate(model.node_ids):
        c = model.coords[i]
        out.append(f"{nid}, " + ", ".join(repr(float(v)) for v in c[:ndim]))
    for k in np.unique(model.kinds):
        out.append(f"*ELEMENT, TYPE={_TYPE_OF_KIND[int(k)]}, ELSET=ALL")
        for e in np.nonzero(model.kinds == k)[0]:
            ids = model.conn_ids[model.conn_ptr[e]:model.conn_ptr[e + 1]]
            out.append(f"{model.elem_ids[e]}, " + ", ".join(str(int(v)) for v in ids))
    for name, E, nu in model.materials:
        out += [f"*MATERIAL, NAME={name}", "*ELASTIC", f"{E!r}, {nu!r}"]
    for es, mat in model.assignments:
        out += [f"*SOLID SECTION, ELSET={es}, MATERIAL={mat}", repr(float(model.thickness))]
    out += ["*STEP", "*STATIC", "*BOUNDARY"]
    for nodes, dofs in model.bcs:
        d = [int(x) for x in dofs]
        contiguous = d == list(range(d[0], d[0] + len(d)))
        for nid in nodes:
            if contiguous:
                out.append(f"{nid}, {d[0]}, {d[-1]}")
            else:
                out += [f"{nid}, {x}, {x}" for x in d]
    out.append("*CLOAD")
    for nid, dof, val in zip(model.cload_node, model.cload_dof, model.cload_value):
        out.append(f"{nid}, {dof}, {float(val)!r}")
    out.append("*END STEP")
    return "\n".join(out) + "\n"
