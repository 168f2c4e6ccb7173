"""Writes snapshot_datasets.json: the datasets the reference's output routine
creates in one t####.h5 file, for the flag settings of the two-fluid runs.

The table is read out of the reference's own source text, not retyped:
this script walks the statements of IO_rot.f90's output path --

  output           IO_rot.f90:38-110   (save_coordinates, write_time,
                                        def_varfiles(0) at nout = 0,
                                        save_varfiles(nout))
  save_coordinates IO_rot.f90:182-191
  write_time       IO_rot.f90:169-180
  def_varfiles     IO_rot.f90:231-259
  save_varfiles    IO_rot.f90:284-330
  set_initial_out  IO_rot.f90:362-374  (the names file_m / file_h)

-- with a small interpreter for the statement forms these routines use
(block and one-line IF, DO loops, CALL, array-constructor assignments) and
a whitelisting evaluator for their conditions, and records every dataset
name in creation order for each case.  A name created twice (the reference
ignores the HDF5 error of the second h5dcreate_f) appears twice.

Usage: python tests/golden/make_snapshot_table.py [IO_rot.f90]
(default: /root/reference/IO_rot.f90).  The output is committed; the tests
read only the JSON.
"""
from __future__ import annotations

import ast
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_SRC = "/root/reference/IO_rot.f90"

# flags the conditions read that the two-fluid runs hold fixed: flag_mhd = 1,
# flag_pip = 1, no ambipolar term, resistivity, gravity or viscosity
FIXED = dict(flag_mhd=1, flag_pip=1, flag_amb=0, flag_resi=0, flag_grav=0, flag_visc=0, flag_rad=0,
             nvar_h=5, et_sav=0, gr_sav=0, vs_sav=0)
SAV = ("ac_sav", "xi_sav", "col_sav", "ion_sav", "rec_sav", "heat_sav", "ps_sav")

CASES = [
    # name, flags, nout; flag_col = 1 (constant alpha_c) is what the runs use
    ("brio_wu_1d", dict(nvar_m=8, flag_divb=0, flag_col=1, flag_ir=0, flag_ir_type=0), 0),
    ("orszag_tang", dict(nvar_m=9, flag_divb=1, flag_col=1, flag_ir=0, flag_ir_type=0), 0),
    ("orszag_tang_later_output", dict(nvar_m=9, flag_divb=1, flag_col=1, flag_ir=0, flag_ir_type=0), 1),
    ("orszag_tang_ps_sav", dict(nvar_m=9, flag_divb=1, flag_col=1, flag_ir=0, flag_ir_type=0, ps_sav=1), 0),
    ("col_sav", dict(nvar_m=9, flag_divb=1, flag_col=1, flag_ir=0, flag_ir_type=0, col_sav=1), 1),
    ("collisional_col_2", dict(nvar_m=9, flag_divb=1, flag_col=2, flag_ir=0, flag_ir_type=0), 0),
    ("ir_1_type_1", dict(nvar_m=9, flag_divb=1, flag_col=1, flag_ir=1, flag_ir_type=1), 0),
    ("kelvin_helmholtz_ir_3_type_0", dict(nvar_m=9, flag_divb=1, flag_col=1, flag_ir=3, flag_ir_type=0), 0),
    ("kelvin_helmholtz_ir_3_type_0_later", dict(nvar_m=9, flag_divb=1, flag_col=1, flag_ir=3,
                                                flag_ir_type=0), 5),
    ("ir_3_type_0_heat_ion_sav", dict(nvar_m=9, flag_divb=1, flag_col=1, flag_ir=3, flag_ir_type=0,
                                      heat_sav=1, ion_sav=1), 0),
    ("ir_3_type_0_col_3_later", dict(nvar_m=9, flag_divb=1, flag_col=3, flag_ir=3, flag_ir_type=0), 1),
    ("ac_xi_sav", dict(nvar_m=9, flag_divb=1, flag_col=1, flag_ir=0, flag_ir_type=0, ac_sav=1, xi_sav=1), 0),
]

_F2PY = [(r"\.eq\.", " == "), (r"\.ne\.", " != "), (r"\.ge\.", " >= "), (r"\.gt\.", " > "),
         (r"\.le\.", " <= "), (r"\.lt\.", " < "), (r"\.and\.", " and "), (r"\.or\.", " or "),
         (r"\.not\.", " not ")]
_OK = (ast.Expression, ast.BoolOp, ast.And, ast.Or, ast.UnaryOp, ast.Not, ast.Compare, ast.Eq, ast.NotEq,
       ast.Gt, ast.GtE, ast.Lt, ast.LtE, ast.Name, ast.Load, ast.Constant)


def cond(text: str, env: dict) -> bool:
    """A Fortran logical expression over integer variables and literals."""
    s = text.lower()
    for a, b in _F2PY:
        s = re.sub(a, b, s)
    tree = ast.parse(s.strip(), mode="eval")
    for node in ast.walk(tree):
        if not isinstance(node, _OK):
            raise ValueError(f"unsupported condition form {text!r}")
        if isinstance(node, ast.Constant) and not isinstance(node.value, int):
            raise ValueError(f"non-integer literal in {text!r}")

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.BoolOp):
            vals = [ev(v) for v in n.values]
            return all(vals) if isinstance(n.op, ast.And) else any(vals)
        if isinstance(n, ast.UnaryOp):
            return not ev(n.operand)
        if isinstance(n, ast.Compare):
            left = ev(n.left)
            for op, r in zip(n.ops, n.comparators):
                right = ev(r)
                ok = {ast.Eq: left == right, ast.NotEq: left != right, ast.Gt: left > right,
                      ast.GtE: left >= right, ast.Lt: left < right, ast.LtE: left <= right}[type(op)]
                if not ok:
                    return False
                left = right
            return True
        if isinstance(n, ast.Name):
            if n.id not in env:
                raise KeyError(f"condition reads {n.id}, which the case does not set")
            return env[n.id]
        return n.value
    return bool(ev(tree))


def statements(src: str, name: str):
    """The statements of subroutine `name`, lower-cased, comments and
    continuations removed."""
    lines = src.lower().splitlines()
    start = next(i for i, l in enumerate(lines) if re.match(rf"\s*subroutine\s+{name}\b", l))
    end = next(i for i in range(start, len(lines)) if re.match(rf"\s*end\s+subroutine\s+{name}\b", lines[i]))
    out, buf = [], ""
    for l in lines[start + 1:end]:
        l = re.sub(r"!.*$", "", l).replace("\t", " ").strip()
        if not l:
            continue
        if l.endswith("&"):
            buf += l[:-1] + " "
            continue
        out.append((buf + l.lstrip("&")).strip())
        buf = ""
    return out


def paren_end(s: str, i: int) -> int:
    """Index just past the parenthesis group opening at s[i]."""
    depth = 0
    for k in range(i, len(s)):
        depth += {"(": 1, ")": -1}.get(s[k], 0)
        if depth == 0:
            return k + 1
    raise ValueError(f"unbalanced parentheses in {s!r}")


def split_args(s: str):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
            continue
        depth += {"(": 1, "[": 1, ")": -1, "]": -1}.get(ch, 0)
        cur += ch
    out.append(cur.strip())
    return out


class Walker:
    def __init__(self, src: str):
        self.src = src
        self.created: list[str] = []

    def name_of(self, expr: str, env: dict) -> str:
        m = re.fullmatch(r"[\"'](\w+)[\"']", expr)
        if m:
            return m.group(1)
        m = re.fullmatch(r"trim\((file_[mh])\((\w+)\)\)", expr)
        if m:
            arr, idx = env[m.group(1)], m.group(2)
            k = env[idx] if not idx.isdigit() else int(idx)
            return arr[k]
        raise ValueError(f"unsupported dataset name expression {expr!r}")

    def run(self, name: str, env: dict, args=()):
        body = statements(self.src, name)
        self.block(body, 0, len(body), env)

    def block(self, st, i, end, env):
        while i < end:
            s = st[i]
            m = re.match(r"if\s*\(", s)
            if m:
                j = paren_end(s, s.index("("))
                c, rest = s[s.index("(") + 1:j - 1], s[j:].strip()
                if rest == "then":
                    # block IF: find its ENDIF (no ELSE in these routines)
                    depth, k = 1, i + 1
                    while depth:
                        t = st[k]
                        if re.match(r"if\s*\(.*\)\s*then$", t):
                            depth += 1
                        elif re.fullmatch(r"end\s*if", t):
                            depth -= 1
                        elif t.startswith("else"):
                            raise ValueError("ELSE is not handled")
                        k += 1
                    if cond(c, env):
                        self.block(st, i + 1, k - 1, env)
                    i = k
                    continue
                if cond(c, env):
                    self.simple(rest, env)
                i += 1
                continue
            m = re.fullmatch(r"do\s+(\w+)\s*=\s*(\w+)\s*,\s*(\w+)(\+1)?", s)
            if m:
                depth, k = 1, i + 1
                while depth:
                    if re.match(r"do\s+\w+\s*=", st[k]):
                        depth += 1
                    elif re.fullmatch(r"end\s*do", st[k]):
                        depth -= 1
                    k += 1
                lo = int(m.group(2)) if m.group(2).isdigit() else env[m.group(2)]
                hi = int(m.group(3)) if m.group(3).isdigit() else env[m.group(3)]
                hi += 1 if m.group(4) else 0
                for v in range(lo, hi + 1):
                    self.block(st, i + 1, k - 1, {**env, m.group(1): v})
                i = k
                continue
            self.simple(s, env)
            i += 1

    def simple(self, s, env):
        m = re.match(r"call\s+(\w+)\s*(\((.*)\))?$", s)
        if m:
            callee, args = m.group(1), split_args(m.group(3) or "")
            if callee == "write_3d_array":
                self.created.append(self.name_of(args[0], env))
            elif callee == "write_1d_array":
                self.created.append(self.name_of(args[1], env))
            elif callee == "h5dcreate_f":
                self.created.append(self.name_of(args[1], env))
            elif callee in ("def_varfiles", "save_coordinates", "write_time", "save_varfiles"):
                sub = {**env}
                if callee == "save_varfiles":
                    sub["n_out"] = env["nout"] if args[0] == "nout" else int(args[0])
                self.run(callee, sub)
            # every other CALL (dataspaces, file handling, IOT) creates no dataset
            return
        m = re.fullmatch(r"(file_[mh])\((\d+)(?::(\d+))?\)\s*=\s*(.+)", s)
        if m:
            arr = env.setdefault(m.group(1), {})
            lo = int(m.group(2))
            hi = int(m.group(3) or lo)
            vals = re.findall(r"'(\w+)'", m.group(4))
            assert len(vals) == hi - lo + 1, s
            for k, v in zip(range(lo, hi + 1), vals):
                arr[k] = v
            return
        if re.match(r"(allocate|integer|character|double|write\s*\(6|dims_|offset|per\s*=|varname)", s):
            return
        raise ValueError(f"unhandled statement {s!r}")


def case_table(src: str, flags: dict, nout: int):
    env = {**FIXED, **{k: 0 for k in SAV}, **flags, "nout": nout}
    w = Walker(src)
    w.run("set_initial_out", env)
    # output (IO_rot.f90:91-96): coordinates, time, def_varfiles(0) at nout = 0,
    # save_varfiles(nout); the calls are read from its statements
    for s in statements(src, "output"):
        for callee in ("save_coordinates", "write_time", "def_varfiles", "save_varfiles"):
            m = re.match(rf"(if\s*\((.*)\)\s*)?call\s+{callee}\b\s*(\((.*)\))?$", s)
            if not m:
                continue
            if m.group(2) and not cond(m.group(2), env):
                continue
            w.simple(s[m.start(0) + (len(m.group(1)) if m.group(1) else 0):], env)
    return w.created


def main():
    src_path = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_SRC
    src = open(src_path).read()
    cases = []
    for name, flags, nout in CASES:
        full = {**{k: 0 for k in SAV}, **flags}
        cases.append(dict(name=name, flags=full, nout=nout, datasets=case_table(src, flags, nout)))
    out = dict(source="IO_rot.f90: output 38-110, write_time 169-180, save_coordinates 182-191, "
                      "def_varfiles 231-259, save_varfiles 284-330, set_initial_out 362-374",
               fixed_flags=FIXED, cases=cases)
    path = os.path.join(HERE, "snapshot_datasets.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    for c in cases:
        print(c["name"], c["nout"], c["datasets"])


if __name__ == "__main__":
    main()
