"""Generates pip_adj_gen.h: reverse-mode (vector-Jacobian) code for the
face fluxes and the coupling source of the two-fluid step.

The forward computations are restated here as SSA statements (one
assignment per intermediate; mathematically the reference's formulas, with
shared reciprocals and primitive-variable forms where they save divisions)
organised in a tree of branches that mirrors the reference (HLL_rot.f90:615-687 HLLC, :691-910 HLLD with the HLLC-G fallback
and the GLM psi flux, PIP_rot.f90:1359-1368 collision source).  For every
statement y = f(x1..xk) the generator emits the local chain rule
x_i_bar += y_bar * df/dx_i (derivatives by sympy), in reverse order, after
re-running the forward statements of the taken branch.  Branch decisions
(wave-speed ordering, HLLD region, min/max, floors) are those of the forward
point: the result is the exact derivative of the piecewise-smooth map at
that point.

Run:  python pip_amd/csrc/gen_adjoint.py  (rewrites pip_adj_gen.h)
"""
from __future__ import annotations

import os

import sympy as sp
from sympy.printing.c import C99CodePrinter

HERE = os.path.dirname(os.path.abspath(__file__))


class Printer(C99CodePrinter):
    def _print_Pow(self, e):
        b, x = e.base, e.exp
        pb = self.parenthesize(b, 100)
        if x == 2:
            return f"({pb}*{pb})"
        if x == 3:
            return f"({pb}*{pb}*{pb})"
        if x == -1:
            return f"(1.0/{pb})"
        if x == -2:
            return f"(1.0/({pb}*{pb}))"
        if x == sp.Rational(1, 2):
            return f"sqrt({self._print(b)})"
        if x == sp.Rational(-1, 2):
            return f"(1.0/sqrt({self._print(b)}))"
        if x == sp.Rational(-3, 2):
            return f"(1.0/({pb}*sqrt({self._print(b)})))"
        return super()._print_Pow(e)

    def _print_Rational(self, e):
        return f"({float(e.p)}/{float(e.q)})"

    def _print_Integer(self, e):
        return f"{float(e)}"

    def _print_Float(self, e):
        return repr(float(e))

    def _print_sign(self, e):
        a = self._print(e.args[0])
        return f"(({a}) > 0.0 ? 1.0 : (({a}) < 0.0 ? -1.0 : 0.0))"


PR = Printer()


def cc(e) -> str:
    return PR.doprint(e)


class Block:
    """A straight-line sequence of statements plus optional branches."""

    def __init__(self, g):
        self.g = g
        self.items = []  # ('let', sym, expr, kind) | ('if', [(cond, Block), ...])

    def let(self, name, expr, kind="expr", exact=False):
        """exact: the statement and the temporaries it lowers to are evaluated
        without contraction (a sum of products that cancels bit for bit in
        the rounded forward arithmetic stays exactly zero)."""
        self.g._exact_mode = exact
        e = self.g.lower(self, sp.sympify(expr)) if kind == "expr" else sp.sympify(expr)
        self.g._exact_mode = False
        s = self.g.sym(name)
        if exact:
            self.g.exact.add(s)
        self.items.append(("let", s, e, kind))
        return s

    def out(self, name, expr):
        """Assign an output symbol (may be assigned in several branches)."""
        e = self.g.lower(self, sp.sympify(expr))
        s = self.g.syms.get(name) or self.g.sym(name)
        self.items.append(("let", s, e, "expr"))
        return s

    def minv(self, name, a, b):
        return self.let(name, sp.Min(a, b, evaluate=False), "min")

    def maxv(self, name, a, b):
        return self.let(name, sp.Max(a, b, evaluate=False), "max")

    def select(self, name, cond, a, b):
        """name = cond ? a : b for symbols/constants a, b; the reverse routes
        the adjoint to the taken operand without a branch."""
        s = self.g.sym(name)
        self.items.append(("let", s, (cond, sp.sympify(a), sp.sympify(b)), "sel"))
        return s

    def switch(self, name, x, eps):
        """name = (|x| + x) / max(2 x, eps) for a symbol x (HLL_rot.f90:679-681):
        1 for 2 x > eps, 0 for x <= 0, 2 x / eps in between.  Its derivative
        is 2 / eps inside that band and exactly 0 outside; the general rule
        for the quotient would form the zero as the difference of two terms
        of size 1 / x, whose rounding leaves a residual that grows as x -> 0."""
        s = self.g.sym(name)
        self.items.append(("let", s, (sp.sympify(x), float(eps)), "switch"))
        return s

    def const(self, name, code):
        """A value with zero derivative (e.g. sign), given as C code."""
        s = self.g.sym(name)
        self.items.append(("let", s, code, "const"))
        return s

    def branches(self, *conds):
        """conds: C conditions; the last may be None (else).  Returns blocks."""
        blocks = [Block(self.g) for _ in conds]
        self.items.append(("if", list(zip(conds, blocks))))
        return blocks


class Gen:
    def __init__(self, name):
        self.name = name
        self.syms = {}
        self.inputs = []
        self.exact = set()  # statements evaluated without contraction
        self._exact_mode = False
        self.root = Block(self)

    def sym(self, name):
        if name in self.syms:
            raise ValueError("redefined " + name)
        s = sp.Symbol(name, real=True)
        self.syms[name] = s
        return s

    def inp(self, name):
        s = self.sym(name)
        self.inputs.append(s)
        return s

    # ------------------------------------------------------------ lowering
    # Every quotient and square root becomes its own statement, so that its
    # reverse costs one division (q = t_bar / d; n_bar += q; d_bar -= q t)
    # instead of symbolic derivatives that divide again by d or d^2; the
    # remaining expressions are polynomials in statement symbols.
    def _temp(self, blk, expr, kind):
        self._ntemp = getattr(self, "_ntemp", 0) + 1
        t = self.sym(f"{self.name}{self._ntemp}")
        if self._exact_mode:
            self.exact.add(t)
        blk.items.append(("let", t, expr, kind))
        return t

    def _atom(self, blk, e):
        return e if (e.is_Symbol or e.is_Number) else self._temp(blk, e, "expr")

    def lower(self, blk, e):
        if e.is_Atom:
            return e
        if isinstance(e, sp.Pow):
            b, x = e.args
            if x == sp.Rational(1, 2):
                return self._temp(blk, self._atom(blk, self.lower(blk, b)), "sqrt")
            if x.is_Rational and x < 0:
                d = self.lower(blk, sp.Pow(b, -x))
                return self._temp(blk, (sp.Integer(1), self._atom(blk, d)), "div")
            return sp.Pow(self.lower(blk, b), x)
        if isinstance(e, sp.Mul):
            num, den = [], []
            for f in e.args:
                if isinstance(f, sp.Pow) and f.args[1].is_Rational and f.args[1] < 0:
                    den.append(sp.Pow(f.args[0], -f.args[1]))
                else:
                    num.append(f)
            n = sp.Mul(*[self.lower(blk, f) for f in num])
            if not den:
                return n
            d = self.lower(blk, sp.Mul(*den))
            return self._temp(blk, (self._atom(blk, n), self._atom(blk, d)), "div")
        return e.func(*[self.lower(blk, a) for a in e.args])

    # ------------------------------------------------------------ emission
    def _fwd(self, blk, out, ind):
        pad = " " * ind
        for it in blk.items:
            if it[0] == "let":
                _, s, e, kind = it
                if kind == "const":
                    out.append(f"{pad}{s} = {e};")
                elif kind == "min":
                    a, b = e.args
                    out.append(f"{pad}{s} = ({cc(a)} < {cc(b)}) ? {cc(a)} : {cc(b)};")
                elif kind == "max":
                    a, b = e.args
                    out.append(f"{pad}{s} = ({cc(a)} > {cc(b)}) ? {cc(a)} : {cc(b)};")
                elif kind == "sel":
                    out.append(f"{pad}{s} = ({e[0]}) ? {cc(e[1])} : {cc(e[2])};")
                elif kind == "switch":
                    x, eps = cc(e[0]), e[1]
                    out.append(f"{pad}{s} = (2.0 * {x} > {eps!r}) ? 1.0 : (({x} > 0.0) ? {x} * {2.0 / eps!r} : 0.0);")
                elif kind == "div":
                    out.append(f"{pad}{s} = PIP_VDIV({cc(e[0])}, {cc(e[1])});")
                elif kind == "sqrt":  # constants fold through sqrt
                    out.append(f"{pad}{s} = {'sqrt' if e.is_Number else 'PIP_VSQRT'}({cc(e)});")
                elif s in self.exact:
                    out.append(f"{pad}{{ PIP_FWD_EXACT {s} = {cc(e)}; }}")
                else:
                    out.append(f"{pad}{s} = {cc(e)};")
            else:
                pass  # branches handled by _emit

    def _rev_stmts(self, blk, out, ind):
        pad = " " * ind
        for it in reversed(blk.items):
            if it[0] != "let":
                continue
            _, s, e, kind = it
            if kind == "const":
                continue
            ab = f"a_{s}"
            if kind == "sel":
                known = set(self.syms.values())
                cond, x, y = e
                if x in known:
                    out.append(f"{pad}a_{x} += ({cond}) ? {ab} : 0.0;")
                if y in known:
                    out.append(f"{pad}a_{y} += ({cond}) ? 0.0 : {ab};")
                continue
            if kind == "switch":
                x, eps = e
                if x in set(self.syms.values()):
                    out.append(f"{pad}a_{x} += ({cc(x)} > 0.0 && 2.0 * {cc(x)} <= {eps!r}) ? {ab} * {2.0 / eps!r} : 0.0;")
                continue
            if kind == "div":
                n, d = e
                known = set(self.syms.values())
                out.append(f"{pad}{{ const double q_ = PIP_VRDIV({ab}, {cc(d)});")
                if n in known:
                    out.append(f"{pad}  a_{n} += q_;")
                if d in known:
                    out.append(f"{pad}  a_{d} -= q_ * {s}; }}")
                else:
                    out.append(f"{pad}}}")
                continue
            if kind == "sqrt":
                if e in set(self.syms.values()):
                    out.append(f"{pad}a_{e} += PIP_VRDIV(0.5 * {ab}, {s});")
                continue
            if kind in ("min", "max"):
                a, b = e.args
                op = "<" if kind == "min" else ">"
                ta = f"a_{a}" if isinstance(a, sp.Symbol) and a in self.syms.values() else None
                tb = f"a_{b}" if isinstance(b, sp.Symbol) and b in self.syms.values() else None
                out.append(f"{pad}if ({cc(a)} {op} {cc(b)}) {{ {ta + ' += ' + ab + ';' if ta else ''} }}"
                           f" else {{ {tb + ' += ' + ab + ';' if tb else ''} }}")
                continue
            for x in sorted(e.free_symbols, key=lambda q: q.name):
                if x not in self.syms.values():
                    continue
                d = sp.diff(e, x)
                if d == 0:
                    continue
                out.append(f"{pad}a_{x} += {ab} * ({cc(d)});")

    def _emit(self, blk, out, ind, recompute=False):
        """forward of blk, then its branches (each fully emitted), then the
        reverse of blk's own statements.  With recompute, blk's forward is
        evaluated again (from inputs passed through an empty asm barrier, so
        the compiler cannot reuse the first values) before its reverse: its
        intermediates need not stay live across the branches."""
        pad = " " * ind
        # forward statements before the first branch, branches, reverse
        self._fwd(blk, out, ind)
        has_branch = False
        for it in blk.items:
            if it[0] == "if":
                has_branch = True
                first = True
                for cond, child in it[1]:
                    if cond is None:
                        out.append(f"{pad}}} else {{")
                    elif first:
                        out.append(f"{pad}if ({cond}) {{")
                    else:
                        out.append(f"{pad}}} else if ({cond}) {{")
                    first = False
                    self._emit(child, out, ind + 2)
                out.append(f"{pad}}}")
        if recompute and has_branch:
            out.append(f"{pad}// recompute the side quantities for their reverse")
            for x in self.inputs:
                out.append(f"{pad}PIP_OPAQUE({x});")
            self._fwd(blk, out, ind)
        self._rev_stmts(blk, out, ind)

    def all_syms(self):
        return list(self.syms.values())

    def emit(self, signature, prologue, epilogue, out_seeds, recompute_root=False):
        """out_seeds: list of (symbol name, C expression of its adjoint)."""
        body = []
        inter = [s for s in self.all_syms() if s not in self.inputs]
        body.append("  // forward intermediates")
        for i in range(0, len(inter), 8):
            body.append("  double " + ", ".join(str(s) for s in inter[i:i + 8]) + ";")
        body.append("  // adjoints")
        allv = self.all_syms()
        for i in range(0, len(allv), 6):
            body.append("  double " + ", ".join(f"a_{s} = 0.0" for s in allv[i:i + 6]) + ";")
        for name, expr in out_seeds:
            body.append(f"  a_{name} = {expr};")
        self._emit(self.root, body, 2, recompute=recompute_root)
        return signature + " {\n  PIP_VJP_CONTRACT\n" + prologue + "\n".join(body) + "\n" + epilogue + "}\n"


# ====================================================================== HLLC
def gen_hllc():
    g = Gen("hllc")
    vl = [g.inp(f"vl{k}") for k in range(5)]
    vr = [g.inp(f"vr{k}") for k in range(5)]
    gm = sp.Symbol("gm", real=True)
    ig1 = sp.Symbol("ig1", real=True)  # 1/(gm-1), computed once in the prologue
    B = g.root
    acc = sp.Float(1.03)
    epsi = 1.0e-10

    # The conserved states and physical fluxes in primitive form: u1/u0 = v1
    # etc. hold exactly in real arithmetic, so the map (and its derivative)
    # is the reference's, evaluated with fewer divisions.
    def side(v, t):
        u0 = B.let(f"u{t}0", v[0])
        u1 = B.let(f"u{t}1", v[0] * v[1])
        u2 = B.let(f"u{t}2", v[0] * v[2])
        u3 = B.let(f"u{t}3", v[0] * v[3])
        u4 = B.let(f"u{t}4", v[4] * ig1 + sp.Rational(1, 2) * v[0] * (v[1] ** 2 + v[2] ** 2 + v[3] ** 2))
        u = [u0, u1, u2, u3, u4]
        p = v[4]
        f = [B.let(f"f{t}0", u1), B.let(f"f{t}1", p + u1 * v[1]), B.let(f"f{t}2", u1 * v[2]),
             B.let(f"f{t}3", u1 * v[3]), B.let(f"f{t}4", v[1] * (u4 + p))]
        cs = B.let(f"cs{t}", sp.sqrt(gm * p / u0))
        vn = v[1]
        return u, f, cs, vn, p

    ul, fl, csl, vnl, pl = side(vl, "l")
    ur, fr, csr, vnr, pr = side(vr, "r")
    a1 = B.let("sla", vnl - csl)
    a2 = B.let("slb", vnr - csr)
    mn = B.minv("slm", a1, a2)
    sl = B.let("sl", acc * mn)
    b1 = B.let("sra", vnl + csl)
    b2 = B.let("srb", vnr + csr)
    mx = B.maxv("srm", b1, b2)
    sr = B.let("sr", acc * mx)
    # exact: mirror-image states (a mirror wall) give S_M = 0 bit for bit
    sm = B.let("sm", (ur[1] * (sr - vnr) - ul[1] * (sl - vnl) + pl - pr) / (ur[0] * (sr - vnr) - ul[0] * (sl - vnl)),
               exact=True)
    prm = B.let("prm", sm * ul[0] * (sl - vnl) - ul[1] * (sl - vnl) + pl)
    il = B.let("il", 1 / (sl - sm))
    ir = B.let("ir", 1 / (sr - sm))
    stl0 = B.let("stl0", ul[0] * (sl - vnl) * il)
    str0 = B.let("str0", ur[0] * (sr - vnr) * ir)
    stl = [stl0, B.let("stl1", stl0 * sm), B.let("stl2", stl0 * vl[2]), B.let("stl3", stl0 * vl[3]),
           B.let("stl4", ((sl - vnl) * ul[4] - pl * vnl + prm * sm) * il)]
    st_r = [str0, B.let("str1", str0 * sm), B.let("str2", str0 * vr[2]), B.let("str3", str0 * vr[3]),
            B.let("str4", ((sr - vnr) * ur[4] - pr * vnr + prm * sm) * ir)]
    # smooth switches (HLL_rot.f90:679-681): derivative inside the epsi band
    # is kept; outside it they are constants 0/1
    thl = B.switch("thl", sl, epsi)
    thm = B.switch("thm", sm, epsi)
    thr = B.let("thr", 1 - B.switch("swr", sr, epsi))
    for k in range(5):
        B.let(f"F{k}", thr * fr[k] + thl * fl[k] + (1 - thl - thr) *
              (thm * (fl[k] + sl * (stl[k] - ul[k])) + (1 - thm) * (fr[k] + sr * (st_r[k] - ur[k]))))
    sig = "PIP_HD void hllc_flux_vjp(const double* vl, const double* vr, double gm, const double* lam, double* lvl, double* lvr)"
    pro = "".join(f"  const double vl{k} = vl[{k}], vr{k} = vr[{k}];\n" for k in range(5)) + \
        "  const double ig1 = 1.0 / (gm - 1.0);\n"
    epi = "".join(f"  lvl[{k}] += a_vl{k};\n  lvr[{k}] += a_vr{k};\n" for k in range(5))
    seeds = [(f"F{k}", f"lam[{k}]") for k in range(5)]
    return g.emit(sig, pro, epi, seeds)


# ====================================================================== HLLD
def gen_hlld():
    g = Gen("hlld")
    NV = 9
    vl = [g.inp(f"vl{k}") for k in range(NV)]
    vr = [g.inp(f"vr{k}") for k in range(NV)]
    gm = sp.Symbol("gm", real=True)
    ch = sp.Symbol("ch", real=True)
    ig1 = sp.Symbol("ig1", real=True)  # 1/(gm-1), computed once in the prologue
    h = sp.Float(1.001)
    R = g.root
    half = sp.Rational(1, 2)

    def side(v, t):
        b2 = R.let(f"b2{t}", v[5] ** 2 + v[6] ** 2 + v[7] ** 2)
        u = [v[0], R.let(f"u{t}1", v[0] * v[1]), R.let(f"u{t}2", v[0] * v[2]), R.let(f"u{t}3", v[0] * v[3]),
             R.let(f"u{t}4", v[4] * ig1 + half * v[0] * (v[1] ** 2 + v[2] ** 2 + v[3] ** 2) + half * b2),
             v[5], v[6], v[7], v[8]]
        pt = R.let(f"pt{t}", v[4] + half * b2)
        vb = R.let(f"vb{t}", v[1] * v[5] + v[2] * v[6] + v[3] * v[7])
        f = [R.let(f"f{t}0", v[0] * v[1]),
             R.let(f"f{t}1", v[0] * v[1] ** 2 + v[4] + half * b2 - v[5] ** 2),
             R.let(f"f{t}2", v[0] * v[1] * v[2] - v[5] * v[6]),
             R.let(f"f{t}3", v[0] * v[1] * v[3] - v[5] * v[7]),
             R.let(f"f{t}4", (u[4] + pt) * v[1] - vb * v[5]),
             sp.Integer(0),
             R.let(f"f{t}6", v[1] * v[6] - v[5] * v[2]),
             R.let(f"f{t}7", v[1] * v[7] - v[5] * v[3])]
        f1 = R.let(f"gp{t}", gm * v[4])
        A = R.let(f"A{t}", f1 + b2)
        Dr = R.let(f"Dr{t}", A ** 2 - 4 * f1 * v[5] ** 2)
        D = R.maxv(f"D{t}", Dr, sp.Float(0.0))
        vf = R.let(f"vf{t}", sp.sqrt((A + sp.sqrt(D)) / (2 * v[0])))
        return u, f, pt, vb, vf

    ul, fl, ptl, vbl, vfl = side(vl, "l")
    ur, fr, ptr, vbr, vfr = side(vr, "r")
    mnv = R.minv("mnv", vl[1], vr[1])
    mxv = R.maxv("mxv", vl[1], vr[1])
    mxf = R.maxv("mxf", vfl, vfr)
    sl = R.let("sl", mnv - mxf)
    sr = R.let("sr", mxv + mxf)

    def glm(blk, outs):
        # flux components 5 and 8 overridden by the GLM terms (HLL_rot.f90:904-908)
        outs[5] = half * (vl[8] + vr[8] - ch * (vr[5] - vl[5]))
        outs.append(half * (ch * ch * (vl[5] + vr[5]) - ch * (vr[8] - vl[8])))
        for k in range(9):
            blk.out(f"F{k}", outs[k])

    bA, bB, bC = R.branches("sl >= 0.0", "sr <= 0.0", None)
    glm(bA, list(fl))
    glm(bB, list(fr))
    inv = bC.let("inv", 1 / (sr - sl))
    # exact: mirror-image states give uh1 = 0, so S_M = 0, bit for bit
    uh = [bC.let(f"uh{k}", inv * (sr * ur[k] - sl * ul[k] - fr[k] + fl[k]), exact=True) for k in range(8)]
    # shared reciprocals: the star states divide by (s - sm) and by uh0
    iu0 = bC.let("iu0", 1 / uh[0])
    sm = bC.let("sm", uh[1] * iu0)
    pt = bC.let("pt", ptl + vl[0] * (sl - vl[1]) * (sm - vl[1]))
    isl = bC.let("isl", 1 / (sl - sm))
    isr = bC.let("isr", 1 / (sr - sm))
    rol1 = bC.let("rol1", vl[0] * (sl - vl[1]) * isl)
    ror1 = bC.let("ror1", vr[0] * (sr - vr[1]) * isr)
    rls = bC.let("rls", sp.sqrt(rol1))
    rrs = bC.let("rrs", sp.sqrt(ror1))
    sal = bC.let("sal", sp.Abs(uh[5]) / rls)
    sar = bC.let("sar", sp.Abs(uh[5]) / rrs)
    cG, cD = bC.branches("(sl >= (sm - 1.001 * sal)) || ((sm + 1.001 * sar) >= sr)", None)
    # ---- HLLC-G fallback
    vy2 = cG.let("gvy2", uh[2] * iu0)
    vz2 = cG.let("gvz2", uh[3] * iu0)
    gL, gR = cG.branches("sm >= 0.0", None)
    for blk, v, u, f, s, ro1, ptk, vb, t, isd in ((gL, vl, ul, fl, sl, rol1, ptl, vbl, "L", isl),
                                                  (gR, vr, ur, fr, sr, ror1, ptr, vbr, "R", isr)):
        u24 = blk.let(f"g{t}u24", ((s - v[1]) * u[4] - ptk * v[1] + pt * sm +
                                   uh[5] * (vb - sm * uh[5] - vy2 * uh[6] - vz2 * uh[7])) * isd)
        u2 = [None, ro1 * sm, ro1 * vy2, ro1 * vz2, u24]
        outs = [s * (ro1 - v[0]) + f[0]] + [s * (u2[k] - u[k]) + f[k] for k in range(1, 5)] + \
               [s * (uh[k] - v[k]) + f[k] for k in range(5, 8)]
        glm(blk, outs)
    # ---- HLLD
    bh2 = cD.let("bh2", uh[5] ** 2)

    def star(v, u, s, ro1, ptk, vb, t, isd):
        f1 = cD.let(f"d{t}f1", 1 / (v[0] * (s - v[1]) * (s - sm) - bh2))
        e = cD.let(f"d{t}e", v[0] * (s - v[1]) ** 2 - bh2)
        u16 = cD.let(f"d{t}u16", v[6] * f1 * e)
        u17 = cD.let(f"d{t}u17", v[7] * f1 * e)
        vy1 = cD.let(f"d{t}vy1", v[2] - f1 * uh[5] * v[6] * (sm - v[1]))
        vz1 = cD.let(f"d{t}vz1", v[3] - f1 * uh[5] * v[7] * (sm - v[1]))
        u14 = cD.let(f"d{t}u14", ((s - v[1]) * u[4] - ptk * v[1] + pt * sm +
                                  uh[5] * (vb - sm * uh[5] - vy1 * u16 - vz1 * u17)) * isd)
        u1 = [ro1, ro1 * sm, ro1 * vy1, ro1 * vz1, u14, uh[5], u16, u17]
        return u1, vy1, vz1

    ul1, vyl1, vzl1 = star(vl, ul, sl, rol1, ptl, vbl, "L", isl)
    ur1, vyr1, vzr1 = star(vr, ur, sr, ror1, ptr, vbr, "R", isr)
    sl1 = cD.let("sl1", sm - sal)
    sr1 = cD.let("sr1", sm + sar)
    dL, dR, dC = cD.branches("sl1 >= 0.0", "sr1 <= 0.0", None)
    for blk, v, u, f, s, u1 in ((dL, vl, ul, fl, sl, ul1), (dR, vr, ur, fr, sr, ur1)):
        outs = [s * (u1[0] - v[0]) + f[0]] + [s * (u1[k] - u[k]) + f[k] for k in range(1, 5)] + \
               [s * (u1[k] - v[k]) + f[k] for k in range(5, 8)]
        glm(blk, outs)
    fc = dC.let("fc", 1 / (rls + rrs))
    sg = dC.const("sg", "copysign(1.0, uh5)")
    cvy2 = dC.let("cvy2", fc * (rls * vyl1 + rrs * vyr1 + (ur1[6] - ul1[6]) * sg))
    cvz2 = dC.let("cvz2", fc * (rls * vzl1 + rrs * vzr1 + (ur1[7] - ul1[7]) * sg))
    u26 = dC.let("u26", fc * (rls * ur1[6] + rrs * ul1[6] + rls * rrs * (vyr1 - vyl1) * sg))
    u27 = dC.let("u27", fc * (rls * ur1[7] + rrs * ul1[7] + rls * rrs * (vzr1 - vzl1) * sg))
    # L** (sm >= 0) and R** fluxes share one code path over the side
    # operands selected by the sign of sm: adjacent faces of a smooth flow
    # take both signs, and separate branches would run both per wavefront.
    cond = "sm >= 0.0"

    def pick(name, a, b):
        return dC.select(name, cond, a, b)
    v = {k: pick(f"cv{k}", vl[k], vr[k]) for k in (0, 5, 6, 7)}
    u = {k: pick(f"cu{k}", ul[k], ur[k]) for k in range(1, 5)}
    f = {k: pick(f"cf{k}", fl[k], fr[k]) for k in (0, 1, 2, 3, 4, 6, 7)}
    f[5] = sp.Integer(0)  # overridden by the GLM term
    s_ = pick("cs", sl, sr)
    s1 = pick("cs1", sl1, sr1)
    ro1 = pick("cro1", rol1, ror1)
    vy1 = pick("cvy1", vyl1, vyr1)
    vz1 = pick("cvz1", vzl1, vzr1)
    u14 = pick("cu14", ul1[4], ur1[4])
    u16 = pick("cu16", ul1[6], ur1[6])
    u17 = pick("cu17", ul1[7], ur1[7])
    ros = pick("cros", rls, rrs)
    sgn = dC.const("csgn", "((sm >= 0.0) ? -1.0 : 1.0)")
    u1 = [ro1, ro1 * sm, ro1 * vy1, ro1 * vz1, u14, uh[5], u16, u17]
    u24 = dC.let("cu24", u1[4] + sgn * ros * (vy1 * u1[6] + vz1 * u1[7] - cvy2 * u26 - cvz2 * u27) * sg)
    u2 = [u1[0], u1[0] * sm, u1[0] * cvy2, u1[0] * cvz2, u24, uh[5], u26, u27]
    outs = [s1 * u2[0] - (s1 - s_) * u1[0] - s_ * v[0] + f[0]] + \
           [s1 * u2[k] - (s1 - s_) * u1[k] - s_ * u[k] + f[k] for k in range(1, 5)] + \
           [s1 * u2[k] - (s1 - s_) * u1[k] - s_ * v[k] + f[k] for k in range(5, 8)]
    glm(dC, outs)
    sig = ("PIP_HD void hlld9_flux_vjp(const double* vl, const double* vr, double gm, double ch, "
           "const double* lam, double* lvl, double* lvr)")
    pro = "".join(f"  double vl{k} = vl[{k}], vr{k} = vr[{k}];\n" for k in range(9)) + \
        "  const double ig1 = 1.0 / (gm - 1.0);\n"
    epi = "".join(f"  lvl[{k}] += a_vl{k};\n  lvr[{k}] += a_vr{k};\n" for k in range(9))
    seeds = [(f"F{k}", f"lam[{k}]") for k in range(9)]
    return g.emit(sig, pro, epi, seeds, recompute_root=RECOMPUTE)


# ================================================================= source
def gen_source():
    """Collision source (explicit, PIP_rot.f90:1359-1368) as a function of the
    conserved states and alpha; the cq2pv rewrites are identities of the
    state (floors inactive).  Outputs S_m[0..4]; S_h = -S_m."""
    g = Gen("src")
    um = [g.inp(f"um{k}") for k in range(8)]
    uh = [g.inp(f"uh{k}") for k in range(5)]
    ac = g.inp("ac")
    gm = sp.Symbol("gm", real=True)
    B = g.root
    half = sp.Rational(1, 2)
    nde = uh[0]
    nvx, nvy, nvz = B.let("nvx", uh[1] / nde), B.let("nvy", uh[2] / nde), B.let("nvz", uh[3] / nde)
    npr = B.let("npr", (gm - 1) * (uh[4] - half * nde * (nvx ** 2 + nvy ** 2 + nvz ** 2)))
    de = um[0]
    vx, vy, vz = B.let("vx", um[1] / de), B.let("vy", um[2] / de), B.let("vz", um[3] / de)
    pr = B.let("pr", (gm - 1) * (um[4] - half * de * (vx ** 2 + vy ** 2 + vz ** 2) -
                                half * (um[5] ** 2 + um[6] ** 2 + um[7] ** 2)))
    te = B.let("te", pr / de * gm * half)
    nte = B.let("nte", npr / nde * gm)
    B.let("S1", ac * (um[0] * uh[1] - uh[0] * um[1]))
    B.let("S2", ac * (um[0] * uh[2] - uh[0] * um[2]))
    B.let("S3", ac * (um[0] * uh[3] - uh[0] * um[3]))
    B.let("S4", ac * nde * de * (half * ((nvx ** 2 - vx ** 2) + (nvy ** 2 - vy ** 2) + (nvz ** 2 - vz ** 2)) +
                                 sp.Rational(3, 2) / gm * (nte - te)))
    sig = ("PIP_HD void source_vjp(const double* umv, const double* uhv, double acv, double gm, "
           "const double* lam, double* lum, double* luh, double* lac)")
    pro = "".join(f"  const double um{k} = umv[{k}];\n" for k in range(8)) + \
        "".join(f"  const double uh{k} = uhv[{k}];\n" for k in range(5)) + "  const double ac = acv;\n"
    epi = "".join(f"  lum[{k}] += a_um{k};\n" for k in range(8)) + \
        "".join(f"  luh[{k}] += a_uh{k};\n" for k in range(5)) + "  *lac += a_ac;\n"
    seeds = [(f"S{k}", f"lam[{k}]") for k in range(1, 5)]
    return g.emit(sig, pro, epi, seeds)


RECOMPUTE = os.environ.get("PIP_ADJ_RECOMPUTE", "1") == "1"


def main(path=None):
    parts = [
        "// pip_adj_gen.h -- GENERATED by gen_adjoint.py; do not edit by hand.\n"
        "// Vector-Jacobian products of the HLLC flux (HLL_rot.f90:615-687), the\n"
        "// HLLD flux with HLLC-G fallback and GLM psi terms (HLL_rot.f90:691-910)\n"
        "// and the collision source (PIP_rot.f90:1359-1368), at the branch taken\n"
        "// by the forward point.  Each function accumulates into its outputs.\n"
        "#pragma once\n#include \"pip_phys.h\"\n\n"
        "// Empty asm barrier: the compiler must treat x as a new value, so a\n"
        "// recomputation from it is not merged with the first evaluation.\n"
        "#if defined(__HIP_DEVICE_COMPILE__)\n"
        "#define PIP_OPAQUE(x) __asm__ volatile(\"\" : \"+v\"(x))\n"
        "#else\n"
        "#define PIP_OPAQUE(x) __asm__ volatile(\"\" : \"+x\"(x))\n"
        "#endif\n\n"
        "// Quotients of the VJPs.  The adjoint needs derivatives to ~1e-10, not the\n"
        "// forward's correctly rounded quotients, so on the device a quotient is the\n"
        "// hardware reciprocal refined by two Newton-Raphson steps\n"
        "// (r <- r + r (1 - d r)) times n, with one residual correction\n"
        "// (q <- q + r (n - d q)) in the restated forward (PIP_VDIV: within about an\n"
        "// ulp of n / d) and without it in the reverse sweep (PIP_VRDIV: within a\n"
        "// few ulp); square roots take the general sequence without its range\n"
        "// scaling (qsqrt).  The VJP bodies also let the compiler contract\n"
        "// multiply-adds.  The host build divides exactly.\n"
        "#if defined(__HIP_DEVICE_COMPILE__)\n"
        "__device__ __forceinline__ double pip_vrcp(double d) {\n"
        "  double r = __builtin_amdgcn_rcp(d);\n"
        "  r = __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);\n"
        "  return __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);\n"
        "}\n"
        "__device__ __forceinline__ double pip_vdiv(double n, double d) {\n"
        "  const double r = pip_vrcp(d);\n"
        "  const double q = n * r;\n"
        "  return __builtin_fma(r, __builtin_fma(-d, q, n), q);\n"
        "}\n"
        "#define PIP_VDIV(n, d) pip_vdiv((n), (d))\n"
        "#define PIP_VRDIV(n, d) ((n) * pip_vrcp(d))\n"
        "#define PIP_VJP_CONTRACT _Pragma(\"clang fp contract(fast)\")\n"
        "#define PIP_FWD_EXACT _Pragma(\"clang fp contract(off)\")\n"
        "#else\n"
        "#define PIP_VDIV(n, d) ((n) / (d))\n"
        "#define PIP_VRDIV(n, d) ((n) / (d))\n"
        "#define PIP_VJP_CONTRACT\n"
        "#define PIP_FWD_EXACT\n"
        "#endif\n"
        "#define PIP_VSQRT(x) qsqrt(x)\n\n"
        "namespace pip {\n\n",
        gen_hllc(), "\n", gen_hlld(), "\n", gen_source(), "\n}  // namespace pip\n",
    ]
    path = path or os.path.join(HERE, "pip_adj_gen.h")
    with open(path, "w") as f:
        f.write("".join(parts))
    print("wrote", path)


if __name__ == "__main__":
    main()
