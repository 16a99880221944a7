"""Generate ``rules_data.h``: the node/weight tables that define the discretization.

Two families of numbers have to be identical to the reference for the library to be a
drop-in (the caller hands us ``srcvals`` sampled at these nodes, and the near weights
are expressed on them):

* Vioreanu-Rokhlin (RV) nodes and weights on the simplex (0,0),(1,0),(0,1), orders
  0..20 -- the reference keeps them as tables in ``src/tria_routs/koorn-uvs-dat.txt``
  and ``koorn-wts-dat.txt`` (GPL-2-or-later, Z. Gimbutas), read by
  ``get_vioreanu_nodes``/``get_vioreanu_wts`` (``src/tria_routs/koornexps.f90:57-154``).
* Xiao-Gimbutas D3-symmetric triangle rules, degrees 1..30 -- the reference stores the
  orbit generators in ``src/tria_routs/triasymq.f`` (GPL-2-or-later, H. Xiao and
  Z. Gimbutas) and expands them by symmetry (``quaequad0``/``quaenodes``/``triasimp``/
  ``trianmap``, ``triasymq.f:89-230,6559-6894``).

This script reads those two files once (they exist only in the build container) and
writes the numbers, already expanded to the standard simplex, into a C header of our
own layout.  The header is committed; nothing at run time reads the reference.
Run:  python fmm3dbie_amd/csrc/gen_rules_data.py /root/reference
"""
import math
import re
import sys
from pathlib import Path

MAX_RV = 20
MAX_XG = 30


def _num(s):
    return float(s.strip().replace("D", "e").replace("d", "e"))


def parse_rv(ref):
    d = Path(ref) / "src" / "tria_routs"
    uv_txt = (d / "koorn-uvs-dat.txt").read_text().splitlines()
    w_txt = (d / "koorn-wts-dat.txt").read_text().splitlines()
    uvs = {}
    wts = {}
    cur = None
    for line in uv_txt:
        m = re.match(r"\s*if\(norder\.eq\.\s*(\d+)\)", line)
        if m:
            cur = int(m.group(1))
            uvs[cur] = {}
            continue
        m = re.match(r"\s*uvs\((\d),\s*(\d+)\)\s*=(.*)", line)
        if m and cur is not None:
            uvs[cur][(int(m.group(1)), int(m.group(2)))] = _num(m.group(3))
    cur = None
    for line in w_txt:
        m = re.match(r"\s*if\(norder\.eq\.\s*(\d+)\)", line)
        if m:
            cur = int(m.group(1))
            wts[cur] = {}
            continue
        m = re.match(r"\s*wts\(\s*(\d+)\)\s*=(.*)", line)
        if m and cur is not None:
            wts[cur][int(m.group(1))] = _num(m.group(2))
    out = {}
    for n in range(MAX_RV + 1):
        p = (n + 1) * (n + 2) // 2
        u = [uvs[n][(1, i)] for i in range(1, p + 1)]
        v = [uvs[n][(2, i)] for i in range(1, p + 1)]
        w = [wts[n][i] for i in range(1, p + 1)]
        out[n] = (u, v, w)
    return out


def parse_xg_generators(ref):
    """Return {degree: [(x, y, w), ...]} as stored (equilateral-triangle frame)."""
    txt = (Path(ref) / "src" / "tria_routs" / "triasymq.f").read_text()
    gens = {}
    for deg in range(1, MAX_XG + 1):
        arrs = []
        for name in ("xs", "ys", "ws"):
            m = re.search(r"[Dd]ata\s+%s%d/(.*?)/" % (name, deg), txt, re.S)
            body = m.group(1)
            vals = []
            for line in body.splitlines():
                line = line.strip()
                if not line:
                    continue
                # drop the fixed-form continuation marker (first token)
                parts = line.split()
                tok = parts[-1].rstrip(",")
                vals.append(_num(tok))
            arrs.append(vals)
        gens[deg] = list(zip(*arrs))
    return gens


# --- D3 symmetric expansion: follows triasymq.f:6379-6555 (normalisation into the
# lower-left sixth), :6758-6859 (orbit expansion), :204-230 (map to the simplex) and
# :161-200 (weight rescaling so the weights sum to the simplex area 1/2).

_S3 = math.sqrt(3.0)


def _rot(x, y):
    pi = math.atan(1.0) * 4
    th = 2 * pi / 3
    a11 = math.cos(th)
    a22 = math.cos(th)
    a12 = -math.sin(th)
    a21 = -a12
    return a11 * x + a12 * y, a21 * x + a22 * y


def _inside(t, x, y):
    if t == 2:
        ok = (x <= 0) and (x >= -1.0)
        if not ((y >= (-1.0 / _S3 - 1.0e-30)) and (y <= x / _S3)):
            ok = False
        return ok
    if t == 1:
        ok = (x <= 0) and (x >= -1.0) and (y >= -1.0 / _S3) and (y <= x / _S3)
        if (x <= 1.0) and (x >= 0) and (y >= -1.0 / _S3) and (y <= -x / _S3):
            ok = True
        return ok
    ok = True
    if y < -1.0 / _S3:
        ok = False
    if y > _S3 * x + 2.0 / _S3:
        ok = False
    if y > -_S3 * x + 2.0 / _S3:
        ok = False
    return ok


def _normalise(x, y):
    for _ in range(10):
        if _inside(2, x, y):
            return x, y
        if _inside(1, x, y):
            return -x, y
        if _inside(0, x, y):
            x, y = _rot(x, y)
            continue
        raise ValueError("generator outside the triangle")
    raise ValueError("normalisation did not converge")


def _expand(gens):
    xs, ys, ws = [], [], []
    eps = 1e-12
    for (x, y, w) in gens:
        if x * x + y * y < eps:
            xs.append(x); ys.append(y); ws.append(w)
            continue
        if (x * x < eps) or (abs(y - x / _S3) < math.sqrt(eps)):
            w0 = w / 3
            x1, y1 = _rot(x, y)
            x2, y2 = _rot(x1, y1)
            xs += [x, x1, x2]; ys += [y, y1, y2]; ws += [w0, w0, w0]
            continue
        w0 = w / 6
        x1, y1 = _rot(x, y)
        x2, y2 = _rot(x1, y1)
        x3, y3 = _rot(-x, y)
        x4, y4 = _rot(x3, y3)
        xs += [x, x1, x2, -x, x3, x4]
        ys += [y, y1, y2, y, y3, y4]
        ws += [w0] * 6
    return xs, ys, ws


def xg_rule_simplex(gens):
    norm = [(*_normalise(x, y), w) for (x, y, w) in gens]
    xs, ys, ws = _expand(norm)
    area = 0.5
    scale = 1.0 / sum(ws) * area
    us, vs, wo = [], [], []
    sc = 1 / math.sqrt(3.0)
    for x, y, w in zip(xs, ys, ws):
        u = 0.5 * (x + 1) - 0.5 * sc * (y + sc)
        v = 1.0 * sc * (y + sc)
        # trianmap onto (0,0),(1,0),(0,1): x = u, y = v (identity affine map)
        us.append(u); vs.append(v); wo.append(w * scale)
    return us, vs, wo


def _fmt(vals):
    return ",\n".join("  %.17e" % v for v in vals)


def main(ref):
    rv = parse_rv(ref)
    gens = parse_xg_generators(ref)
    lines = []
    lines.append("// Generated by gen_rules_data.py -- do not edit.")
    lines.append("// Numeric tables only.  Provenance and licence:")
    lines.append("//  * RV nodes/weights: Vioreanu-Rokhlin discretization nodes on the simplex,")
    lines.append("//    tables by Z. Gimbutas, GPL-2-or-later (reference src/tria_routs/koorn-*-dat.txt).")
    lines.append("//  * XG rules: Xiao-Gimbutas D3-symmetric triangle quadratures, GPL-2-or-later")
    lines.append("//    (reference src/tria_routs/triasymq.f), expanded here to the simplex.")
    lines.append("#pragma once")
    lines.append("#define RULES_RV_MAXORDER %d" % MAX_RV)
    lines.append("#define RULES_XG_MAXDEG %d" % MAX_XG)
    # RV: packed arrays with offsets
    offs = [0]
    u_all, v_all, w_all = [], [], []
    for n in range(MAX_RV + 1):
        u, v, w = rv[n]
        u_all += u; v_all += v; w_all += w
        offs.append(len(u_all))
    lines.append("static const int rules_rv_off[%d] = {%s};" % (len(offs), ",".join(map(str, offs))))
    lines.append("static const double rules_rv_u[%d] = {\n%s};" % (len(u_all), _fmt(u_all)))
    lines.append("static const double rules_rv_v[%d] = {\n%s};" % (len(v_all), _fmt(v_all)))
    lines.append("static const double rules_rv_w[%d] = {\n%s};" % (len(w_all), _fmt(w_all)))
    offs = [0, 0]
    u_all, v_all, w_all = [], [], []
    for d in range(1, MAX_XG + 1):
        u, v, w = xg_rule_simplex(gens[d])
        u_all += u; v_all += v; w_all += w
        offs.append(len(u_all))
    lines.append("static const int rules_xg_off[%d] = {%s};" % (len(offs), ",".join(map(str, offs))))
    lines.append("static const double rules_xg_u[%d] = {\n%s};" % (len(u_all), _fmt(u_all)))
    lines.append("static const double rules_xg_v[%d] = {\n%s};" % (len(v_all), _fmt(v_all)))
    lines.append("static const double rules_xg_w[%d] = {\n%s};" % (len(w_all), _fmt(w_all)))
    out = Path(__file__).with_name("rules_data.h")
    out.write_text("\n".join(lines) + "\n")
    print("wrote", out, "rv pts", offs[-1])


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
