"""Generate the coefficient tables of the narrow-stencil SBP second derivative with a variable
coefficient, laplace_sbp42_narrow.h and laplace_sbp63_narrow.h, from the reference's
differential_operators/horizontal/laplace/laplace_Ah_sbp{42,63}_narrow_mod.f90
(calc_laplace_tile).

The reference writes each closure row r of D2(b) f as
    d2(r) = scale_r * sum_c m_rc f(c),   m_rc = sum_k (+-) coef_k * b(k)
with b(k) = Qi11 J (x) or Qi22 J (y) at the k-th point from the panel edge; scale_r is
-1/h_rr (sbp42) or 1 (sbp63), coef_k a rational p/q written in place (sbp42) or a named
PARAMETER constant (sbp63).  The interior row has 2R+1 coefficients c(-R..R), each a signed
sum of coefficients times b(i-R..i+R).  This script reads those expressions, checks that the
four closures (x/y, left/right mirror) carry identical ones, and writes every coefficient as a
C constant expression in the reference's operation order, so that C and HIP code fold it to
the same double as the Fortran compiler and then evaluate each product coef*b and the sums
left to right.

Run once, with the reference checked out:  python3 gen_laplace_tables.py /path/to/reference
The generated headers are committed; nothing reads the reference at build or run time.
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "src/differential_operators/horizontal/laplace/laplace_Ah_sbp{o}_narrow_mod.f90"
NUM = r"(\d+(?:\.\d*)?)(?:_8)?"


def fortran_to_c(expr):
    """A Fortran constant expression of real(8) literals as a C expression (same operations)."""
    e = re.sub(r"(\d+\.\d*)_8", r"\1", expr.strip())
    e = re.sub(r"(\d+)_8", r"\1", e)
    if not re.fullmatch(r"[0-9.+\-*/() ]+", e):
        raise ValueError(expr)
    return "(" + e + ")"


def joined_lines(text):
    out, cur = [], ""
    for raw in text.splitlines():
        line = raw.split("!")[0].rstrip()
        if line.endswith("&"):
            cur += line[:-1] + " "
            continue
        out.append(cur + line)
        cur = ""
    return out


def terms(expr, params):
    """Signed (coefficient C expression, b name) terms of a coefficient expression."""
    pat = re.compile(r"([+-])?\s*(?:" + NUM + r"\s*/\s*" + NUM + r"|([A-Za-z]\w*))\s*\*\s*(b\w+)")
    expr = expr.strip()
    res, pos = [], 0
    for m in pat.finditer(expr):
        gap = expr[pos:m.start()].strip()
        if gap:
            raise ValueError(f"unparsed text {gap!r} in {expr!r}")
        sign = -1 if m.group(1) == "-" else 1
        if m.group(4):
            coef = params[m.group(4)]
        else:
            coef = fortran_to_c(f"{m.group(2)}/{m.group(3)}")
        res.append((sign, coef, m.group(5)))
        pos = m.end()
    if expr[pos:].strip():
        raise ValueError(f"unparsed tail in {expr!r}")
    return res


def parse(path):
    text = open(path).read()
    params = {}
    head = text[:text.index("contains")]
    for ln in joined_lines(head):
        for m in re.finditer(r"\b(d\d{6}|di_\w+)\s*=\s*([^,]+?(?:\([^)]*\)[^,]*)?)\s*(?:,|$)", ln):
            params[m.group(1)] = fortran_to_c(m.group(2))
    body = text[text.index("subroutine calc_laplace_tile"):text.index("end subroutine calc_laplace_tile")]
    lines = joined_lines(body)
    blocks, cur = [], None
    for ln in lines:
        s = ln.strip()
        if s.startswith("if (mesh_xy%"):
            cur = {"b": {}, "m": {}, "h": None}
            continue
        if cur is None:
            continue
        mb = re.fullmatch(r"(b\d+)\s*=\s*mesh_xy%Qi\((\d),([^,]+),([^,]+),k\)\*mesh_xy%J\(([^,]+),([^,]+),k\)", s)
        if mb:
            cur["b"][mb.group(1)] = (int(mb.group(2)), mb.group(3).strip(), mb.group(4).strip())
            continue
        mh = re.fullmatch(r"h(\d)(\d)\s*=\s*(.+)", s)
        if mh:
            cur["h"] = fortran_to_c(mh.group(3))
            continue
        mm = re.fullmatch(r"m(\d)(\d+)\s*=\s*(.+)", s)
        if mm:
            cur["m"][int(mm.group(2))] = terms(mm.group(3), params)
            cur["row"] = int(mm.group(1))
            continue
        md = re.fullmatch(r"d2f_([xy])%p\(([^=]+)\)\s*=\s*(-1\.0_8/h\d\d|1\.0_8)\*\((.+)\)", s)
        if md and cur.get("m"):
            cur["dir"], cur["target"] = md.group(1), md.group(2)
            cur["scale"] = "neg_inv_h" if md.group(3).startswith("-") else "one"
            cur["cols"] = re.findall(r"m\d(\d+)\*f%p\(([^)]*)\)", md.group(4))
            assert [int(c) for c, _ in cur["cols"]] == list(range(1, len(cur["cols"]) + 1))
            blocks.append(cur)
            cur = None
    return params, blocks, lines


def generate(order, ref):
    params, blocks, lines = parse(os.path.join(ref, SRC.format(o=order)))
    nb = {42: 6, 63: 9}[order]
    assert len(blocks) == 4 * nb, len(blocks)
    canon = {}
    for b in blocks:
        d = b["dir"]
        idx = b["target"].split(",")[0 if d == "x" else 1].strip()
        right = "n" in idx
        row = b["row"]
        qi = {"x": 1, "y": 3}[d]
        for name, (q, i1, i2) in b["b"].items():
            k = int(name[1:])
            pos = i1 if d == "x" else i2
            want = (f"mesh_xy%n{d}" + ("" if k == 1 else f"-{k - 1}")) if right else str(k)
            assert q == qi and pos.replace(" ", "") == want, (d, right, name, q, pos)
        for c, arg in b["cols"]:
            pos = arg.split(",")[0 if d == "x" else 1].replace(" ", "")
            c = int(c)
            want = (f"mesh_xy%n{d}" + ("" if c == 1 else f"-{c - 1}")) if right else str(c)
            assert pos == want, (d, right, row, c, pos)
        key = (b["scale"], b["h"], tuple((c, tuple(b["m"][c])) for c in sorted(b["m"])))
        if row in canon:
            assert canon[row] == key, f"closure row {row} differs between x/y or left/right"
        canon[row] = key
    rows = sorted(canon)
    assert rows == list(range(1, nb + 1))
    scales = {canon[r][0] for r in rows}
    assert len(scales) == 1
    scale = scales.pop()

    R = {42: 2, 63: 3}[order]
    names = {"bmmm": -3, "bmm": -2, "bm": -1, "bc": 0, "bp": 1, "bpp": 2, "bppp": 3}
    cnames = {"cmmm": -3, "cmm": -2, "cm": -1, "cc": 0, "cp": 1, "cpp": 2, "cppp": 3}
    inter = {}
    for ln in lines:
        m = re.fullmatch(r"\s*(cmmm|cmm|cm|cc|cp|cpp|cppp)\s*=\s*(.+)", ln)
        if m:
            t = [(sg, cf, names[bn]) for sg, cf, bn in terms(m.group(2), params)]
            assert inter.setdefault(cnames[m.group(1)], t) == t
    assert sorted(inter) == list(range(-R, R + 1)), sorted(inter)
    pen = set(re.findall(r"penalty = " + NUM + r"/" + NUM + r"\*", "\n".join(lines)))
    assert len(pen) == 1
    pnum, pden = pen.pop()

    ncol = max(max(c for c, _ in canon[r][2]) for r in rows)
    tc, ts, tb, start, ncols = [], [], [], [], []
    for r in rows:
        _, _, ms = canon[r]
        ncols.append(max(c for c, _ in ms))
        row_start = []
        for c in range(1, ncol + 1):
            row_start.append(len(tc))
            for cc, tl in ms:
                if cc != c:
                    continue
                for sg, cf, bn in tl:
                    tc.append(cf); ts.append(str(sg)); tb.append(str(int(bn[1:]) - 1))
        row_start.append(len(tc))
        start.append(row_start)
    ic, isg, ib, istart = [], [], [], []
    for off in range(-R, R + 1):
        istart.append(len(ic))
        for sg, cf, bo in inter[off]:
            ic.append(cf); isg.append(str(sg)); ib.append(str(bo))
    istart.append(len(ic))
    hs = [canon[r][1] or "(1.0)" for r in rows]

    P = f"LN{order}_"

    def arr(vals, per):
        return "\n".join("    " + ", ".join(vals[i:i + per]) + "," for i in range(0, len(vals), per))

    out = os.path.join(HERE, f"laplace_sbp{order}_narrow.h")
    rowval = ("-((1.0 / h) * (m[r][0] f[0] + m[r][1] f[1] + ...))" if scale == "neg_inv_h"
              else "1.0 * (m[r][0] f[0] + m[r][1] f[1] + ...)")
    with open(out, "w") as f:
        f.write(f"""/* laplace_sbp{order}_narrow.h -- GENERATED by gen_laplace_tables.py; do not edit.
 *
 * Narrow-stencil SBP second derivative D2(b) with a variable coefficient, as the reference's
 * laplace_Ah_sbp{order}_narrow_t (differential_operators/horizontal/laplace/
 * laplace_Ah_sbp{order}_narrow_mod.f90, calc_laplace_tile).  Closure row r = 0..{nb - 1} (r-th point
 * from a panel edge, mirrored at the far edge), with h = {P}H[r]:
 *     d2[r] = {rowval}   ({P}ROW_NEG_INV_H = {1 if scale == "neg_inv_h" else 0})
 *     m[r][c] = sum over t in [{P}MSTART[r][c], {P}MSTART[r][c+1]) of
 *               {P}TSIGN[t] * ({P}TCOEF[t] * b[{P}TB[t]]),
 * summed left to right (the first term's sign applied by negation).  b[k], f[c]: the k-th /
 * c-th point from the edge.  Interior row i: d2 = c(-{R}) f[i-{R}] + ... + c({R}) f[i+{R}] with c(o)
 * the terms [{P}ISTART[o+{R}], {P}ISTART[o+{R + 1}]) over b[i + {P}IB[t]].  The boundary penalty of
 * the cross terms uses {P}PC.  Every coefficient is a constant expression in the reference's
 * operation order.
 */
#ifndef LAPLACE_SBP{order}_NARROW_H
#define LAPLACE_SBP{order}_NARROW_H
#ifdef __cplusplus
#define {P}CONST static constexpr
#else
#define {P}CONST static const
#endif
#define {P}NB {nb}
#define {P}NCOL {ncol}
#define {P}R {R}
#define {P}ROW_NEG_INV_H {1 if scale == "neg_inv_h" else 0}
{P}CONST double {P}H[{nb}] = {{{", ".join(hs)}}};
{P}CONST int {P}NCOLS[{nb}] = {{{", ".join(map(str, ncols))}}};
{P}CONST int {P}MSTART[{nb}][{ncol + 1}] = {{
{chr(10).join("    {" + ", ".join(map(str, s)) + "}," for s in start)}
}};
{P}CONST double {P}TCOEF[{len(tc)}] = {{
{arr(tc, 1)}
}};
{P}CONST int {P}TSIGN[{len(ts)}] = {{
{arr(ts, 16)}
}};
{P}CONST int {P}TB[{len(tb)}] = {{
{arr(tb, 16)}
}};
{P}CONST int {P}ISTART[{2 * R + 2}] = {{{", ".join(map(str, istart))}}};
{P}CONST double {P}ICOEF[{len(ic)}] = {{
{arr(ic, 1)}
}};
{P}CONST int {P}ISIGN[{len(isg)}] = {{{", ".join(isg)}}};
{P}CONST int {P}IB[{len(ib)}] = {{{", ".join(ib)}}};
{P}CONST double {P}PC = {fortran_to_c(f"{pnum}/{pden}")};
#endif
""")
    print(f"wrote {out}: {len(tc)} closure terms, {len(ic)} interior terms, scale {scale}")


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    for order in (42, 63):
        generate(order, ref)
