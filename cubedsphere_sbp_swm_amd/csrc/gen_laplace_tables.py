"""Generate laplace_sbp42_narrow.h, the coefficient table of the narrow-stencil SBP second
derivative with a variable coefficient, from the reference's
differential_operators/horizontal/laplace/laplace_Ah_sbp42_narrow_mod.f90 (calc_laplace_tile,
:94-763).

The reference writes each closure row r of D2(b) f as
    d2(r) = -1/h_rr * sum_c m_rc f(c),   m_rc = sum_k (+-) (p_k / q_k) * b(k)
with b(k) = Qi11 J (x) or Qi22 J (y) at the k-th point from the panel edge, and the interior
row as five coefficients c(-2..2), each a signed sum of rationals times b(i-2..i+2).  This
script reads those expressions, checks that the four closures (x/y, left/right mirror) carry
identical ones, and writes the terms in source order so that C and HIP code evaluate them with
the same operations: each product (p/q)*b, the sums left to right.

Run once, with the reference checked out:  python3 gen_laplace_tables.py /path/to/reference
The generated header is committed; nothing reads the reference at build or run time.
"""
import os
import re
import sys

SRC = "src/differential_operators/horizontal/laplace/laplace_Ah_sbp42_narrow_mod.f90"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "laplace_sbp42_narrow.h")

NUM = r"(\d+(?:\.\d*)?)(?:_8)?"
TERM = re.compile(r"([+-])?\s*" + NUM + r"\s*/\s*" + NUM + r"\s*\*\s*(b\w+)")


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


def terms(expr):
    """Signed rational * b-name terms of an expression, in source order."""
    expr = expr.strip()
    res, pos = [], 0
    for m in TERM.finditer(expr):
        gap = expr[pos:m.start()].strip()
        if gap:
            raise ValueError(f"unparsed text {gap!r} in {expr!r}")
        sign = -1 if m.group(1) == "-" else 1
        res.append((sign, m.group(2), m.group(3), m.group(4)))
        pos = m.end()
    if expr[pos:].strip():
        raise ValueError(f"unparsed tail in {expr!r}")
    return res


def rational(expr):
    m = re.fullmatch(r"\s*" + NUM + r"(?:\s*/\s*" + NUM + r")?\s*", expr)
    if not m:
        raise ValueError(expr)
    return m.group(1), m.group(2) or "1"


def lit(s):
    return s if "." in s else s + ".0"


def parse(path):
    text = open(path).read()
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
            cur["h"] = rational(mh.group(3))
            continue
        mm = re.fullmatch(r"m(\d)(\d+)\s*=\s*(.+)", s)
        if mm:
            cur["m"][int(mm.group(2))] = terms(mm.group(3))
            cur["row"] = int(mm.group(1))
            continue
        md = re.fullmatch(r"d2f_([xy])%p\(([^=]+)\)\s*=\s*-1\.0_8/h(\d)(\d)\*\((.+)\)", s)
        if md and cur.get("m"):
            d = md.group(1)
            cols = re.findall(r"m\d(\d+)\*f%p\(([^)]*)\)", md.group(5))
            cur["dir"], cur["target"], cur["cols"] = d, md.group(2), cols
            blocks.append(cur)
            cur = None
    return blocks, lines


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    blocks, lines = parse(os.path.join(ref, SRC))
    assert len(blocks) == 24, len(blocks)  # 6 rows x (x, y) x (left, right)
    canon = {}
    for b in blocks:
        d = b["dir"]
        idx = b["target"].split(",")[0 if d == "x" else 1].strip()
        right = "n" in idx
        row = b["row"]
        # the k-th b is taken k-1 points from the edge, the c-th column of m likewise
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
        key = (b["h"], tuple((c, tuple(b["m"][c])) for c in sorted(b["m"])))
        if row in canon:
            assert canon[row] == key, f"closure row {row} differs between x/y or left/right"
        canon[row] = key
    rows = sorted(canon)
    assert rows == list(range(1, 7))

    # interior coefficients c(-2..2) in terms of b(i-2..i+2)
    names = {"bmm": -2, "bm": -1, "bc": 0, "bp": 1, "bpp": 2}
    inter = {}
    for ln in lines:
        m = re.fullmatch(r"\s*(cmm|cm|cc|cp|cpp)\s*=\s*(.+)", ln)
        if m:
            t = [(sg, p, q, names[bn]) for sg, p, q, bn in terms(m.group(2))]
            off = {"cmm": -2, "cm": -1, "cc": 0, "cp": 1, "cpp": 2}[m.group(1)]
            assert inter.setdefault(off, t) == t
    assert sorted(inter) == [-2, -1, 0, 1, 2]
    pen = set(re.findall(r"penalty = " + NUM + r"/" + NUM + r"\*", "\n".join(lines)))
    assert len(pen) == 1
    pnum, pden = pen.pop()

    tn, tq, ts, tb = [], [], [], []
    start = []
    ncols = []
    for r in rows:
        h, ms = canon[r]
        ncols.append(max(c for c, _ in ms))
        row_start = []
        for c in range(1, 9):
            row_start.append(len(tn))
            for cc, tl in ms:
                if cc != c:
                    continue
                for sg, p, q, bn in tl:
                    tn.append(lit(p)); tq.append(lit(q)); ts.append(str(sg)); tb.append(str(int(bn[1:]) - 1))
        row_start.append(len(tn))
        start.append(row_start)
    it_n, it_q, it_s, it_b, it_start = [], [], [], [], []
    for off in (-2, -1, 0, 1, 2):
        it_start.append(len(it_n))
        for sg, p, q, bo in inter[off]:
            it_n.append(lit(p)); it_q.append(lit(q)); it_s.append(str(sg)); it_b.append(str(bo))
    it_start.append(len(it_n))

    def arr(vals, per=6):
        out = []
        for i in range(0, len(vals), per):
            out.append("    " + ", ".join(vals[i:i + per]) + ",")
        return "\n".join(out)

    hn = [lit(canon[r][0][0]) for r in rows]
    hd = [lit(canon[r][0][1]) for r in rows]
    with open(OUT, "w") as f:
        f.write(f"""/* laplace_sbp42_narrow.h -- GENERATED by gen_laplace_tables.py; do not edit.
 *
 * Narrow-stencil SBP second derivative D2(b) of 4th order interior accuracy, as the reference's
 * laplace_Ah_sbp42_narrow_t (differential_operators/horizontal/laplace/
 * laplace_Ah_sbp42_narrow_mod.f90:94-763).  Closure row r = 0..5 (r-th point from a panel edge,
 * mirrored at the far edge):
 *     d2[r] = (-1 / (LN42_HNUM[r] / LN42_HDEN[r])) * (m[r][0] f[0] + m[r][1] f[1] + ...)
 *     m[r][c] = sum over t in [LN42_MSTART[r][c], LN42_MSTART[r][c+1]) of
 *               LN42_TSIGN[t] * ((LN42_TNUM[t] / LN42_TDEN[t]) * b[LN42_TB[t]]),
 * summed left to right (the first term's sign applied by negation).  b[k], f[c]: the k-th /
 * c-th point from the edge.  Interior row i: d2 = c(-2) f[i-2] + ... + c(2) f[i+2] with c(o)
 * the terms [LN42_ISTART[o+2], LN42_ISTART[o+3]) over b[i + LN42_IB[t]].  The boundary
 * penalty of the cross terms uses LN42_PNUM / LN42_PDEN.
 */
#ifndef LAPLACE_SBP42_NARROW_H
#define LAPLACE_SBP42_NARROW_H
#ifdef __cplusplus
#define LN42_CONST static constexpr
#else
#define LN42_CONST static const
#endif
#define LN42_NB 6
#define LN42_NCOL 8
LN42_CONST double LN42_HNUM[6] = {{{", ".join(hn)}}};
LN42_CONST double LN42_HDEN[6] = {{{", ".join(hd)}}};
LN42_CONST int LN42_NCOLS[6] = {{{", ".join(map(str, ncols))}}};
LN42_CONST int LN42_MSTART[6][9] = {{
{chr(10).join("    {" + ", ".join(map(str, s)) + "}," for s in start)}
}};
LN42_CONST double LN42_TNUM[{len(tn)}] = {{
{arr(tn, 4)}
}};
LN42_CONST double LN42_TDEN[{len(tq)}] = {{
{arr(tq, 4)}
}};
LN42_CONST int LN42_TSIGN[{len(ts)}] = {{
{arr(ts, 16)}
}};
LN42_CONST int LN42_TB[{len(tb)}] = {{
{arr(tb, 16)}
}};
LN42_CONST int LN42_ISTART[6] = {{{", ".join(map(str, it_start))}}};
LN42_CONST double LN42_INUM[{len(it_n)}] = {{{", ".join(it_n)}}};
LN42_CONST double LN42_IDEN[{len(it_q)}] = {{{", ".join(it_q)}}};
LN42_CONST int LN42_ISIGN[{len(it_s)}] = {{{", ".join(it_s)}}};
LN42_CONST int LN42_IB[{len(it_b)}] = {{{", ".join(it_b)}}};
LN42_CONST double LN42_PNUM = {lit(pnum)};
LN42_CONST double LN42_PDEN = {lit(pden)};
#endif
""")
    print(f"wrote {OUT}: {len(tn)} closure terms, {len(it_n)} interior terms")


if __name__ == "__main__":
    main()
