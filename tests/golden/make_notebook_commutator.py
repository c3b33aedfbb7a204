"""Writes notebook_commutator_nf3.json from the reference's golden output.

Scripts/symbolic_hermitians/Symbolic-Hermitian-Commutator.ipynb (cell 20)
prints the nine real components of D = i[H, f] for 3x3 Hermitian H and f, in
the notebook's own symbols (e/u/t flavour labels, _R/_I parts).  This script
reads that printed OUTPUT (not any source code), parses each line into a list
of bilinear terms [coefficient, H component, f component] in this
repository's Hermitian component order, and stores the table as JSON data.

Run in the development container (the reference tree is not present on GPU
machines):  python tests/golden/make_notebook_commutator.py
"""
import json
import os
import re

NB = "/root/reference/Scripts/symbolic_hermitians/Symbolic-Hermitian-Commutator.ipynb"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "notebook_commutator_nf3.json")

FL = {"e": 0, "u": 1, "t": 2}


def comp_name(sym: str) -> str:
    """'Heu_I' -> '01_Im'; 'fee_R' -> '00_Re' (i <= j as in the notebook)."""
    a, b, part = sym[1], sym[2], sym[4]
    i, j = FL[a], FL[b]
    assert i <= j, sym
    return f"{i}{j}_{'Re' if part == 'R' else 'Im'}"


def main():
    nb = json.load(open(NB))
    text = None
    for cell in nb["cells"]:
        for o in cell.get("outputs", []):
            t = "".join(o.get("text", ""))
            if "Dfee_R" in t:
                text = t
    assert text is not None, "golden output not found"
    table = {}
    term_re = re.compile(r"([+-]?)\s*(\d*)\*?(H[eut]{2}_[RI])\*(f[eut]{2}_[RI])")
    for line in text.splitlines():
        line = line.strip()
        if not line:
            continue
        lhs, rhs = line.rstrip(";").split("=")
        lhs = lhs.strip()
        key = comp_name(lhs[1:])
        terms = []
        body = rhs.strip()
        if not body.startswith("-"):
            body = "+" + body
        for m in term_re.finditer(body.replace(" ", "")):
            sign = -1 if m.group(1) == "-" else 1
            coef = int(m.group(2)) if m.group(2) else 1
            terms.append([sign * coef, comp_name(m.group(3)), comp_name(m.group(4))])
        assert terms, line
        table[key] = terms
    assert len(table) == 9
    json.dump({"source": "Symbolic-Hermitian-Commutator.ipynb cell 20 output: D = i[H,f]",
               "terms": table}, open(OUT, "w"), indent=1, sort_keys=True)
    print("wrote", OUT, sum(len(v) for v in table.values()), "terms")


if __name__ == "__main__":
    main()
