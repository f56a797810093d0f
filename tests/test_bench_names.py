"""bench.py resolves every global name its functions use (no GPU needed): a
leg renamed or lost in an edit shows here rather than as a NameError on the
GPU after minutes of timing."""
import ast
import builtins
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_every_global_name_is_defined():
    tree = ast.parse(open(os.path.join(ROOT, "bench.py")).read())
    defined = set(dir(builtins)) | {"__file__", "__name__"}
    for node in tree.body:
        if isinstance(node, (ast.FunctionDef, ast.ClassDef)):
            defined.add(node.name)
        elif isinstance(node, (ast.Import, ast.ImportFrom)):
            defined.update((a.asname or a.name).split(".")[0] for a in node.names)
        elif isinstance(node, ast.Assign):
            for t in node.targets:
                defined.update(n.id for n in ast.walk(t) if isinstance(n, ast.Name))
    missing = set()
    for fn in (n for n in tree.body if isinstance(n, ast.FunctionDef)):
        local = {a.arg for a in ast.walk(fn.args) if isinstance(a, ast.arg)}
        for n in ast.walk(fn):
            if isinstance(n, ast.Name) and isinstance(n.ctx, (ast.Store, ast.Del)):
                local.add(n.id)
            elif isinstance(n, (ast.FunctionDef, ast.Lambda)) and n is not fn:
                if isinstance(n, ast.FunctionDef):
                    local.add(n.name)
                local.update(a.arg for a in ast.walk(n.args) if isinstance(a, ast.arg))
            elif isinstance(n, (ast.Import, ast.ImportFrom)):
                local.update((a.asname or a.name).split(".")[0] for a in n.names)
            elif isinstance(n, ast.comprehension):
                local.update(x.id for x in ast.walk(n.target) if isinstance(x, ast.Name))
            elif isinstance(n, ast.ExceptHandler) and n.name:
                local.add(n.name)
        for n in ast.walk(fn):
            if isinstance(n, ast.Name) and isinstance(n.ctx, ast.Load) and n.id not in local | defined:
                missing.add((fn.name, n.id))
    assert not missing, sorted(missing)


def test_no_code_after_a_return():
    tree = ast.parse(open(os.path.join(ROOT, "bench.py")).read())
    for fn in (n for n in tree.body if isinstance(n, ast.FunctionDef)):
        for st in fn.body[:-1]:
            assert not isinstance(st, ast.Return), f"{fn.name}: statements after its return"
