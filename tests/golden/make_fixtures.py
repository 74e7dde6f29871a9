"""Derive small golden fixtures from the reference's committed example files.

Run in a container that has the ParaFEM sources at /root/reference:
    python tests/golden/make_fixtures.py
Writes (data only, no reference source):
  p121_demo.json        -- control data, neq, total load, published results of
                           examples/5th_ed/p121/demo/p121_demo.res and a
                           SHA-256 of the parsed deck arrays (g_num after
                           ABAQUS2SG, coordinates, rest, loaded nodes) and
                           the loads as printed;
  p121_demo_displ.npz   -- the nodal displacement field of
                           p121_demo.ensi.DISPL-000001 (x, y, z per node, as
                           printed with 4-5 significant digits).
  p124_demo.json        -- control data of examples/5th_ed/p124/demo/p124_demo
                           (.dat, .mat), a SHA-256 of the parsed deck arrays
                           (g_num after ABAQUS2SG, coordinates, rest, element
                           types) and the temperature history its .res prints;
  p124_demo_ttr.npz     -- the nodal temperatures of p124_demo.ensi.NDTTR-*
                           at steps 10, 80 and 150 (as printed, 5 digits).
  p122_demo.json        -- control data and load increments of
                           examples/5th_ed/p122/demo/p122_demo.dat, its
                           restraints (.bnd) and fixed freedoms (.fix), and
                           the per-increment results its .res prints;
  p122_demo_mesh.npz    -- the demo's nodal coordinates and element nodes
                           (.d; a graded mesh no generator here reproduces).
  p121_book.json        -- the generator input of examples/5th_ed/p121/book
                           (p121.mg) and what its p121.res prints (nodes,
                           restrained, equations, iterations, the central
                           displacement, element 1's centroid stresses);
  p123_book.json        -- the generator input of examples/5th_ed/p123/book
                           (p123.mg) and what its p123.res prints (nodes,
                           restrained, equations, iterations, total load,
                           the four potentials).
The xx3_tiny/ directory is a verbatim copy of examples/dev/xx3/demo/xx3-tiny.*
(the deck and its .res/.dis results).
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from parafem_amd import mesh as M  # noqa: E402

REF = "/root/reference/parafem/examples/5th_ed/p121/demo/p121_demo"


def deck_digest(d):
    h = hashlib.sha256()
    for k in ("g_num", "coords", "rest", "load_node"):
        a = np.ascontiguousarray(d[k])
        if a.dtype.kind == "f":
            a = a + 0.0          # -0.0 and 0.0 hash alike
        h.update(k.encode())
        h.update(str(a.dtype).encode())
        h.update(a.tobytes())
    return h.hexdigest()


REF124 = "/root/reference/parafem/examples/5th_ed/p124/demo/p124_demo"
P124_STEPS = (10, 80, 150)


def p124_digest(d):
    h = hashlib.sha256()
    for k in ("g_num", "coords", "rest", "etype"):
        a = np.ascontiguousarray(d[k])
        if a.dtype.kind == "f":
            a = a + 0.0
        h.update(k.encode())
        h.update(str(a.dtype).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def p124_main():
    d = M.read_deck(REF124, "p124")
    res = open(REF124 + ".res").read().splitlines()
    hist = []
    for line in res:
        t = line.split()
        if len(t) in (2, 3) and t[0][0] in "0123456789" and "E" in t[0]:
            hist.append([float(t[0]), float(t[1]), int(t[2]) if len(t) == 3 else None])
    fields = {}
    for step in P124_STEPS:
        lines = open(f"{REF124}.ensi.NDTTR-{step:06d}").read().split("\n")[4:]
        fields[f"step{step}"] = np.array([float(x) for x in lines if x.strip()], dtype=np.float32)
    keys = ("element", "meshgen", "partitioner", "np_types", "nels", "nn", "nr", "nip", "nod",
            "loaded_nodes", "fixed_freedoms", "val0", "dtim", "nstep", "npri", "theta", "tol", "limit", "nres")
    out = {
        "source": "examples/5th_ed/p124/demo/p124_demo.{dat,d,bnd,mat,res,ensi.NDTTR-*}",
        "control": {k: d[k] for k in keys},
        "prop": d["prop"].tolist(),
        "res_processes": int(res[0].split()[4]),
        "res_neq": int(res[1].split()[7]),
        "history": hist,
        "mg": {"nxe": 25, "nye": 25, "nze": 25, "aa": 0.04, "bb": 0.04, "cc": 0.04},
        "deck_sha256": p124_digest(d),
        "ttr_steps": list(P124_STEPS),
    }
    with open(os.path.join(HERE, "p124_demo.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    np.savez_compressed(os.path.join(HERE, "p124_demo_ttr.npz"), **fields)
    print(json.dumps({k: v for k, v in out.items() if k != "history"}, indent=1))


REF122 = "/root/reference/parafem/examples/5th_ed/p122/demo/p122_demo"


def p122_main():
    d = M.read_deck(REF122, "p122")
    res = open(REF122 + ".res").read().splitlines()
    incs = []
    for k, line in enumerate(res):
        if line.startswith("Load Increment"):
            disp = float(res[k + 1].split()[-1])
            sig = [float(t) for t in res[k + 3].split()]
            cj = int(res[k + 4].split()[-1])
            pl = int(res[k + 5].split()[-1])
            incs.append({"increment": int(line.split()[-1]), "displacement": disp, "sigma_zxy": sig,
                         "cj_iterations": cj, "plastic_iterations": pl})
    keys = ("element", "meshgen", "partitioner", "nels", "nn", "nr", "nip", "nod", "fixed_freedoms",
            "loaded_nodes", "phi", "c", "psi", "e", "v", "incs", "plasits", "cjits", "plastol", "cjtol")
    out = {
        "source": "examples/5th_ed/p122/demo/p122_demo.{dat,d,bnd,fix,res}",
        "control": {k: d[k] for k in keys},
        "qinc": d["qinc"].tolist(),
        "rest": d["rest"].tolist(),
        "fix": [[int(n), int(sn), float(v)] for n, sn, v in zip(d["fix_node"], d["fix_sense"], d["fix_val"])],
        "res_processes": int(res[0].split()[4]),
        "res_neq": int(res[1].split()[7]),
        "res_dt": float(res[3].split()[-1]),
        "increments": incs,
    }
    with open(os.path.join(HERE, "p122_demo.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    np.savez_compressed(os.path.join(HERE, "p122_demo_mesh.npz"), coords=d["coords"], g_num=d["g_num"],
                        etype=d["etype"])
    print(json.dumps({k: v for k, v in out.items() if k not in ("rest",)}, indent=1)[:2000])


REF123 = "/root/reference/parafem/examples/5th_ed/p123/book/p123"


def p123_main():
    mg = [line.split() for line in open(REF123 + ".mg").read().splitlines() if line.strip()]
    res = open(REF123 + ".res").read().splitlines()
    k = next(i for i, line in enumerate(res) if line.strip().startswith("Freedom"))
    out = {
        "source": "examples/5th_ed/p123/book/p123.{mg,res}",
        "mg": {"nels": int(mg[1][1]), "nxe": int(mg[1][2]), "nze": int(mg[1][3]), "nip": int(mg[1][4]),
               "aa": float(mg[2][0]), "bb": float(mg[2][1]), "cc": float(mg[2][2]),
               "kx": float(mg[2][3]), "ky": float(mg[2][4]), "kz": float(mg[2][5]),
               "tol": float(mg[3][0]), "limit": int(mg[3][1]),
               "loaded_freedoms": int(mg[4][0]), "fixed_freedoms": int(mg[4][1])},
        "res_processes": int(res[0].split()[4]),
        "res_nn": int(res[1].split()[2]),
        "res_nr": int(res[1].split()[4]),
        "res_neq": int(res[1].split()[7]),
        "res_iters": int(res[3].split()[-1]),
        "res_total_load": res[4].split()[-1],
        "res_potentials": [[int(line.split()[0]), line.split()[1]] for line in res[k + 1:k + 5]],
    }
    with open(os.path.join(HERE, "p123_book.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


REF121B = "/root/reference/parafem/examples/5th_ed/p121/book/p121"


def p121_book_main():
    mg = " ".join(line for line in open(REF121B + ".mg").read().splitlines()[1:]).split()
    res = open(REF121B + ".res").read().splitlines()
    k = next(i for i, line in enumerate(res) if line.startswith("The Centroid point stresses"))
    out = {
        "source": "examples/5th_ed/p121/book/p121.{mg,res}",
        "mg": {"nels": int(mg[1]), "nxe": int(mg[2]), "nze": int(mg[3]), "nod": int(mg[4]), "nip": int(mg[5]),
               "aa": float(mg[6]), "bb": float(mg[7]), "cc": float(mg[8]), "e": float(mg[9]), "v": float(mg[10]),
               "tol": float(mg[11]), "limit": int(mg[12])},
        "res_processes": int(res[0].split()[4]),
        "res_nn": int(res[1].split()[2]),
        "res_nr": int(res[1].split()[4]),
        "res_neq": int(res[1].split()[7]),
        "res_iters": int(next(line for line in res if "iterations to convergence" in line).split()[-1]),
        "res_x1": next(line for line in res if "central nodal displacement" in line).split()[-1],
        "res_sigma_elem1": res[k + 2].split(),
    }
    with open(os.path.join(HERE, "p121_book.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


def main():
    d = M.read_deck(REF)
    res = open(REF + ".res").read().splitlines()
    lines = open(REF + ".ensi.DISPL-000001").read().split("\n")[4:]
    vals = np.array([float(s) for s in lines if s.strip()])
    disp = vals.reshape(3, -1).T.astype(np.float32)
    sig = [float(t) for t in res[10].split()]
    out = {
        "source": "examples/5th_ed/p121/demo/p121_demo.{dat,d,bnd,lds,res,ensi.DISPL-000001}",
        "nels": d["nels"], "nn": d["nn"], "nr": d["nr"], "nod": d["nod"],
        "loaded_nodes": d["loaded_nodes"], "e": d["e"], "v": d["v"], "tol": d["tol"],
        "limit": d["limit"],
        "mg": {"nxe": 20, "nye": 20, "nze": 20, "aa": 0.5, "bb": 0.5, "cc": 0.5},
        "res_processes": int(res[0].split()[4]),
        "res_neq": int(res[1].split()[7]),
        "res_total_load": float(res[4].split()[-1]),
        "res_iters": int(res[5].split()[-1]),
        "res_x1": float(res[7].split()[-1]),
        "res_sigma_elem1": sig,
        "deck_sha256": deck_digest(d),
        "load_node": d["load_node"].tolist(),
        "load_val": d["load_val"].tolist(),
    }
    with open(os.path.join(HERE, "p121_demo.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    np.savez_compressed(os.path.join(HERE, "p121_demo_displ.npz"), disp=disp)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
    p124_main()
    p122_main()
    p123_main()
    p121_book_main()
