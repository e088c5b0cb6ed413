"""Extracts the golden data held in the reference tree into small JSON fixtures.

Run where the reference tree is present (the tests read only the fixtures it writes):
    python tests/golden/make_golden.py
Writes:
  gll_table.json  -- the literal GLL nodes/weights of flat_shell_engine/gll.m:11-34 (n = 2..10)
  m1_rand_2.json  -- the gmsh quad mesh flat_shell/gmsh_out/m1_rand_2.m (POS, QUADS)
The reference files are parsed as text; nothing of the reference is executed.
"""
import json
import re
from pathlib import Path

REF = Path("/root/reference/src/models/flat_shell")
OUT = Path(__file__).resolve().parent


def gll_table():
    txt = (REF / "flat_shell_engine" / "gll.m").read_text()
    table = {}
    for m in re.finditer(r"case\s+(\d+);?(.*?)(?=case|otherwise|end;)", txt, re.S):
        n = int(m.group(1))
        body = m.group(2)
        vals = {}
        for name in ("ksi", "wi"):
            arr = [None] * n
            for a in re.finditer(name + r"\((\d+)\)\s*=\s*([^;]+);", body):
                i = int(a.group(1)) - 1
                expr = a.group(2).strip()
                ref = re.fullmatch(r"(-?)" + name + r"\((\d+)\)", expr)
                if ref:
                    v = arr[int(ref.group(2)) - 1]
                    arr[i] = -v if ref.group(1) else v
                else:
                    arr[i] = float(expr)
            vals[name] = arr
        if n >= 2:
            table[n] = vals
    return table


def m1_rand_2():
    txt = (REF / "gmsh_out" / "m1_rand_2.m").read_text()
    pos = re.search(r"msh.POS = \[(.*?)\];", txt, re.S).group(1)
    quads = re.search(r"msh.QUADS =\[(.*?)\];", txt, re.S).group(1)
    P = [[float(x) for x in ln.replace(";", " ").split()] for ln in pos.strip().splitlines() if ln.strip()]
    Q = [[int(x) for x in ln.split()] for ln in quads.strip().splitlines() if ln.strip()]
    return {"POS": P, "QUADS": Q}


if __name__ == "__main__":
    (OUT / "gll_table.json").write_text(json.dumps(gll_table(), indent=1))
    (OUT / "m1_rand_2.json").write_text(json.dumps(m1_rand_2(), indent=1))
    print("wrote", sorted(p.name for p in OUT.glob("*.json")))
