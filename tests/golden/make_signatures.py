"""Writes reference_signatures.json: the parameter lists (names, kinds and
defaults) of every public function of the reference modules vivsim.lbm,
vivsim.lbm3d, vivsim.ib, vivsim.ib3d, vivsim.dyn, vivsim.post and
vivsim.multigrid.

The reference is read as text with ``ast`` (nothing of it is imported or run;
starred subscripts, Python >= 3.11 syntax, are unstarred first):
each name re-exported by a package ``__init__.py`` is resolved to the module
that defines it, either as a ``def`` or as ``name = wrapper(core)``, in which
case the signature is the one of the wrapper's inner ``def``.  Defaults are
recorded as literals, or as the bare name when they are a name (e.g.
``kernel=kernel_peskin_4pt``).

    python tests/golden/make_signatures.py [/root/reference]
"""

import ast
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _const(node):
    """Value of a literal or of arithmetic on literals (e.g. ``cs2=1.0/3.0``)."""
    if isinstance(node, ast.BinOp):
        a, b = _const(node.left), _const(node.right)
        ops = {ast.Add: lambda: a + b, ast.Sub: lambda: a - b, ast.Mult: lambda: a * b, ast.Div: lambda: a / b,
               ast.Pow: lambda: a ** b}
        return ops[type(node.op)]()
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
        return -_const(node.operand)
    return ast.literal_eval(node)


def _default(node):
    if node is None:
        return None
    if isinstance(node, ast.Name):
        return {"name": node.id}
    return {"value": _const(node)}


def _params(fn):
    a = fn.args
    out = []
    pos = a.posonlyargs + a.args
    defs = [None] * (len(pos) - len(a.defaults)) + list(a.defaults)
    for p, d in zip(pos, defs):
        out.append({"name": p.arg, "kind": "positional", "default": _default(d)})
    if a.vararg:
        out.append({"name": a.vararg.arg, "kind": "var_positional", "default": None})
    for p, d in zip(a.kwonlyargs, a.kw_defaults):
        out.append({"name": p.arg, "kind": "keyword_only", "default": _default(d)})
    if a.kwarg:
        out.append({"name": a.kwarg.arg, "kind": "var_keyword", "default": None})
    return out


def _py310(text):
    """Drops the star of starred subscripts (``f[:, *spec.wall]``, Python >= 3.11
    syntax) so that this interpreter's ast can parse the file; only the
    signatures are read, never the bodies."""
    out, stack, i = [], [], 0
    while i < len(text):
        ch = text[i]
        if ch in "([{":
            stack.append(ch)
        elif ch in ")]}" and stack:
            stack.pop()
        elif ch == "*" and stack and stack[-1] == "[":
            j = len(out) - 1
            while j >= 0 and out[j] in " \t":
                j -= 1
            if j >= 0 and out[j] in ",[" and i + 1 < len(text) and (text[i + 1].isalpha() or text[i + 1] == "_"):
                i += 1
                continue
        out.append(ch)
        i += 1
    return "".join(out)


class Module:
    def __init__(self, path, pkg_root, modname):
        self.path, self.pkg_root, self.modname = path, pkg_root, modname
        self.tree = ast.parse(_py310(open(path).read()))
        self.defs = {n.name: n for n in self.tree.body if isinstance(n, ast.FunctionDef)}
        self.assigns = {}
        self.imports = {}
        for n in self.tree.body:
            if isinstance(n, ast.Assign) and len(n.targets) == 1 and isinstance(n.targets[0], ast.Name):
                self.assigns[n.targets[0].id] = n.value
            if isinstance(n, ast.ImportFrom):
                for al in n.names:
                    self.imports[al.asname or al.name] = (n.module, n.level, al.name)

    def resolve_module(self, module, level):
        if level == 0:
            parts = module.split(".")
            return os.path.join(self.pkg_root, *parts)
        base = os.path.dirname(self.path)
        for _ in range(level - 1):
            base = os.path.dirname(base)
        return os.path.join(base, *(module.split(".") if module else []))


def _load(target, pkg_root):
    path = target + ".py" if os.path.exists(target + ".py") else os.path.join(target, "__init__.py")
    return Module(path, pkg_root, target)


def signature(mod, name, pkg_root, depth=0):
    assert depth < 8, name
    if name in mod.defs:
        return _params(mod.defs[name])
    if name in mod.assigns:
        v = mod.assigns[name]
        if isinstance(v, ast.Call) and isinstance(v.func, ast.Name):
            wrapper = v.func.id
            wmod = mod
            if wrapper in mod.imports:
                m, lvl, orig = mod.imports[wrapper]
                wmod, wrapper = _load(mod.resolve_module(m, lvl), pkg_root), orig
            fn = wmod.defs[wrapper]
            inner = [n for n in fn.body if isinstance(n, ast.FunctionDef)]
            assert inner, (name, wrapper)
            return _params(inner[0])
        if isinstance(v, ast.Name):
            return signature(mod, v.id, pkg_root, depth + 1)
    if name in mod.imports:
        m, lvl, orig = mod.imports[name]
        return signature(_load(mod.resolve_module(m, lvl), pkg_root), orig, pkg_root, depth + 1)
    raise KeyError(f"{name} not found in {mod.path}")


def exported(init_mod):
    names = []
    for n in init_mod.tree.body:
        if isinstance(n, ast.ImportFrom):
            names += [al.asname or al.name for al in n.names]
    return names


def main(ref_root):
    pkg_root = ref_root
    out = {}
    for pkg in ("lbm", "lbm3d", "ib", "ib3d"):
        init = _load(os.path.join(ref_root, "vivsim", pkg), pkg_root)
        out[pkg] = {name: signature(init, name, pkg_root) for name in exported(init)}
    for modname in ("dyn", "post", "multigrid"):
        mod = _load(os.path.join(ref_root, "vivsim", modname), pkg_root)
        out[modname] = {name: _params(fn) for name, fn in mod.defs.items() if not name.startswith("_")}
    dst = os.path.join(HERE, "reference_signatures.json")
    with open(dst, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(f"wrote {dst}: " + ", ".join(f"{k} {len(v)}" for k, v in out.items()))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
